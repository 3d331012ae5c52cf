// C ABI of the batched multi-scene form (skw_batch_*): scene table, work queue (smallest scenes
// first), the frame captured in a hipGraph, and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "internal.hpp"

using namespace skw;

struct skw_batch {
  std::vector<const skw_scene *> scenes;
  std::vector<DevBatchScene> table;
  std::vector<GatherArgs> jobs;   // photon-gather jobs, heap storage attached
  std::vector<void *> job_mem;    // per job: heap slots (nq * k positions, then nq * k d^2)
  GatherArgs *d_jobs = nullptr;
  uint64_t total_rays = 0;
  uint64_t total_queries = 0;
  uint32_t max_stack = 0;          // deepest walk or lookup stack of any item
  int device = 0;
  int cus = 256;
  double wall_khz = 100000.0;
  skw_batch_config cfg{1, 32, 256, 16};
  DevBatchScene *d_table = nullptr;
  DevChunk *d_chunks = nullptr;
  uint64_t n_chunks = 0;
  uint64_t *d_state = nullptr;  // queue head, frame start, stall flag, then per item: done, end
  skw_hit *d_hits = nullptr;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t last_stream = nullptr;

  size_t n_items() const { return table.size() + jobs.size(); }
  BatchArgs args() const {
    BatchArgs a;
    a.scenes = d_table;
    a.jobs = d_jobs;
    a.chunks = d_chunks;
    a.n_chunks = n_chunks;
    a.queue_head = d_state;
    a.frame_start = d_state + 1;
    a.stalled = d_state + 2;
    a.item_done = d_state + 3;
    a.item_end = d_state + 3 + n_items();
    a.n_scenes = (uint32_t)table.size();
    a.n_jobs = (uint32_t)jobs.size();
    a.hits = d_hits;
    return a;
  }
  uint32_t grid() const {
    uint64_t waves = (uint64_t)cus * cfg.waves_per_cu;
    uint64_t blocks = (waves + cfg.waves_per_block - 1) / cfg.waves_per_block;
    return (uint32_t)std::max<uint64_t>(1, blocks);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~skw_batch() {
    drop_graph();
    void *ptrs[] = {d_table, d_chunks, d_state, d_hits, d_jobs};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    for (void *p : job_mem)
      if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

bool valid_config(const skw_batch *b, const skw_batch_config &c) {
  bool waves_ok = c.waves_per_block == 1 || c.waves_per_block == 2 || c.waves_per_block == 4;
  bool stack_ok = (c.stack_entries == 32 || c.stack_entries == 64) && c.stack_entries >= b->max_stack;
  return waves_ok && stack_ok && c.chunk_rays >= 64 && c.chunk_rays % 64 == 0 && c.waves_per_cu >= 1 &&
         c.waves_per_cu <= 64;
}

// Chunks of cfg.chunk_rays rays (or gather queries), items ordered by their size (then index), so
// small scenes and jobs finish first.
hipError_t build_queue(skw_batch *b) {
  const size_t ns = b->table.size();
  auto units = [&](size_t item) { return item < ns ? b->table[item].n_rays : b->jobs[item - ns].nq; };
  std::vector<uint32_t> order(b->n_items());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return units(x) < units(y); });
  std::vector<DevChunk> chunks;
  for (uint32_t s : order) {
    const uint64_t n = units(s);
    for (uint64_t r = 0; r < n; r += b->cfg.chunk_rays) {
      DevChunk c;
      c.item = s;
      c.pad = 0;
      c.begin = r;
      c.end = std::min<uint64_t>(r + b->cfg.chunk_rays, n);
      chunks.push_back(c);
    }
  }
  if (b->d_chunks) (void)hipFree(b->d_chunks);
  b->d_chunks = nullptr;
  b->n_chunks = chunks.size();
  return upload(&b->d_chunks, chunks);
}

hipError_t enqueue_frame(skw_batch *b, hipStream_t st) {
  BatchArgs a = b->args();
  hipError_t e = launch_batch_reset(a, st);
  if (e != hipSuccess) return e;
  return launch_batch_trace(a, b->cfg.waves_per_block, b->cfg.stack_entries, b->grid(), st);
}

hipError_t capture(skw_batch *b) {
  b->drop_graph();
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t e2 = enqueue_frame(b, b->stream);
  e = hipStreamEndCapture(b->stream, &b->graph);
  if (e2 != hipSuccess) return e2;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
}

}  // namespace

extern "C" {

skw_status skw_batch_create(const skw_scene *const *scenes, const skw_batch_view *views, uint32_t n,
                            skw_batch **out) {
  return skw_batch_create_with_gathers(scenes, views, n, nullptr, 0, out);
}

skw_status skw_batch_create_with_gathers(const skw_scene *const *scenes, const skw_batch_view *views, uint32_t n,
                                         const skw_batch_gather *gathers, uint32_t n_gathers, skw_batch **out) {
  if (!out || (n > 0 && (!scenes || !views)) || (n_gathers > 0 && !gathers))
    return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (n == 0 && n_gathers == 0) return fail(SKW_INVALID_ARG, "empty batch");
  *out = nullptr;
  std::unique_ptr<skw_batch> b(new skw_batch());
  SKW_HIP(hipGetDevice(&b->device), "hipGetDevice");
  SKW_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device), "hipDeviceGetAttribute");
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device) == hipSuccess && khz > 0)
    b->wall_khz = khz;
  uint64_t off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const skw_scene *s = scenes[i];
    if (!s || !s->on_device) return fail(SKW_INVALID_ARG, "every scene must be a device scene");
    if (s->has_inst) return fail(SKW_UNSUPPORTED, "batched scenes cannot use instancing");
    if (s->host.stack_need > 64) return fail(SKW_UNSUPPORTED, "a batched scene needs a stack deeper than 64");
    const skw_batch_view &v = views[i];
    if (v.width == 0 || v.height == 0 || v.spp == 0) return fail(SKW_INVALID_ARG, "empty view");
    DevBatchScene t;
    std::memset(&t, 0, sizeof(t));
    t.nodes = s->d_nodes;
    t.tris = s->d_tris;
    t.root = s->host.root;
    t.width = v.width;
    t.height = v.height;
    t.spp = v.spp;
    for (int k = 0; k < 3; ++k) t.cam_o[k] = v.cam.origin[k];
    for (int k = 0; k < 9; ++k) t.cam_r[k] = v.cam.rot[k];
    t.x_tan = v.cam.x_tan;
    t.y_tan = v.cam.y_tan;
    t.seed = v.seed;
    t.n_rays = (uint64_t)v.width * v.height * v.spp;
    t.hit_offset = off;
    t.nested = s->host.nested ? 1u : 0u;
    off += t.n_rays;
    b->table.push_back(t);
    b->scenes.push_back(s);
    b->max_stack = std::max(b->max_stack, s->host.stack_need);
  }
  b->total_rays = off;
  for (uint32_t j = 0; j < n_gathers; ++j) {
    const skw_batch_gather &g = gathers[j];
    if (!g.map || !g.map->on_device) return fail(SKW_INVALID_ARG, "every gather needs a device photon map");
    if (g.nq == 0 || g.k == 0) return fail(SKW_INVALID_ARG, "empty gather job");
    if (!g.d_points || !g.d_normals || !g.d_e_ext || !g.d_e_int)
      return fail(SKW_INVALID_ARG, "gather points, normals, e_ext and e_int must not be NULL");
    if (!(g.r >= 0.0)) return fail(SKW_INVALID_ARG, "radius must be >= 0");
    if (g.map->max_depth + 1 > 64) return fail(SKW_UNSUPPORTED, "a photon tree needs a stack deeper than 64");
    GatherArgs a = gather_args(g.map, g.d_points, g.d_normals, g.nq, g.k, g.r, g.paths, g.rho, 1);
    a.e_ext = g.d_e_ext;
    a.e_int = g.d_e_int;
    a.radiance = g.d_radiance;
    a.defer = 1;  // no ids or d^2 are returned: slots are written only for sets that fill
    const size_t slots = (size_t)g.nq * g.k;
    void *mem = nullptr;
    SKW_HIP(hipMalloc(&mem, slots * (sizeof(uint32_t) + sizeof(double)) + 256), "gather heap allocation");
    b->job_mem.push_back(mem);
    a.heap_d2 = static_cast<double *>(mem);
    a.heap_pos = reinterpret_cast<uint32_t *>(static_cast<char *>(mem) + slots * sizeof(double));
    b->jobs.push_back(a);
    b->total_queries += g.nq;
    b->max_stack = std::max(b->max_stack, a.stack_need);
  }
  b->cfg.stack_entries = b->max_stack <= 32 ? 32 : 64;
  SKW_HIP(upload(&b->d_table, b->table), "upload scene table");
  SKW_HIP(upload(&b->d_jobs, b->jobs), "upload gather jobs");
  SKW_HIP(hipMalloc((void **)&b->d_state, (3 + 2 * b->n_items()) * sizeof(uint64_t)), "state allocation");
  SKW_HIP(hipMalloc((void **)&b->d_hits, std::max<uint64_t>(1, off) * sizeof(skw_hit)), "hit allocation");
  SKW_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream creation");
  SKW_HIP(build_queue(b.get()), "queue upload");
  SKW_HIP(capture(b.get()), "frame capture");
  *out = b.release();
  return SKW_OK;
}

skw_status skw_batch_destroy(skw_batch *batch) {
  if (batch) (void)hipDeviceSynchronize();
  delete batch;
  return SKW_OK;
}

skw_status skw_batch_configure(skw_batch *b, const skw_batch_config *cfg) {
  if (!b || !cfg) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (!valid_config(b, *cfg)) return fail(SKW_INVALID_ARG, "invalid launch shape for this batch");
  SKW_HIP(hipDeviceSynchronize(), "synchronize");
  b->cfg = *cfg;
  SKW_HIP(build_queue(b), "queue upload");
  SKW_HIP(capture(b), "frame capture");
  return SKW_OK;
}

skw_status skw_batch_frame(skw_batch *b, void *stream) {
  if (!b) return fail(SKW_INVALID_ARG, "batch must not be NULL");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SKW_HIP(hipGraphLaunch(b->exec, st), "graph launch");
  b->last_stream = st;
  return SKW_OK;
}

skw_status skw_batch_autotune(skw_batch *b, uint32_t frames, skw_batch_config *best, double *best_ms) {
  if (!b) return fail(SKW_INVALID_ARG, "batch must not be NULL");
  if (frames == 0) frames = 3;
  const uint32_t waves_opts[] = {1, 2, 4};
  const uint32_t stack_opts[] = {32, 64};
  const uint32_t chunk_opts[] = {64, 256, 1024, 4096};
  const uint32_t wpc_opts[] = {8, 16, 32};
  skw_batch_config keep = b->cfg, win = b->cfg;
  double win_ms = 1e300;
  hipEvent_t e0, e1;
  SKW_HIP(hipEventCreate(&e0), "event");
  SKW_HIP(hipEventCreate(&e1), "event");
  for (uint32_t w : waves_opts)
    for (uint32_t s : stack_opts)
      for (uint32_t c : chunk_opts)
        for (uint32_t p : wpc_opts) {
          skw_batch_config cfg{w, s, c, p};
          if (!valid_config(b, cfg)) continue;
          if (s == 64 && b->max_stack <= 32) continue;  // a deeper LDS stack only lowers occupancy
          b->cfg = cfg;
          if (build_queue(b) != hipSuccess) continue;
          hipError_t e = enqueue_frame(b, b->stream);
          e = e == hipSuccess ? hipEventRecord(e0, b->stream) : e;
          for (uint32_t f = 0; f < frames && e == hipSuccess; ++f) e = enqueue_frame(b, b->stream);
          e = e == hipSuccess ? hipEventRecord(e1, b->stream) : e;
          e = e == hipSuccess ? hipEventSynchronize(e1) : e;
          float ms = 0.0f;
          if (e == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
            double per = ms / frames;
            if (per < win_ms) {
              win_ms = per;
              win = cfg;
            }
          }
        }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (win_ms == 1e300) {
    b->cfg = keep;
    (void)build_queue(b);
    (void)capture(b);
    return fail(SKW_HIP_ERROR, "no launch shape ran");
  }
  b->cfg = win;
  SKW_HIP(build_queue(b), "queue upload");
  SKW_HIP(capture(b), "frame capture");
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return SKW_OK;
}

skw_status skw_batch_hits(const skw_batch *b, uint32_t scene, const skw_hit **d_hits, uint64_t *n) {
  if (!b || scene >= b->table.size()) return fail(SKW_INVALID_ARG, "bad batch or scene index");
  if (d_hits) *d_hits = b->d_hits + b->table[scene].hit_offset;
  if (n) *n = b->table[scene].n_rays;
  return SKW_OK;
}

skw_status skw_batch_latency(skw_batch *b, double *ns) {
  if (!b || !ns) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  SKW_HIP(hipDeviceSynchronize(), "synchronize");
  std::vector<uint64_t> st(3 + 2 * b->n_items());
  SKW_HIP(hipMemcpy(st.data(), b->d_state, st.size() * sizeof(uint64_t), hipMemcpyDeviceToHost), "state copy");
  const size_t n = b->n_items();
  for (size_t i = 0; i < n; ++i) {
    uint64_t end = st[3 + n + i];
    ns[i] = end >= st[1] ? (double)(end - st[1]) * 1e6 / b->wall_khz : 0.0;
  }
  if (st[2] != 0) return fail(SKW_HIP_ERROR, "a wave of the last batch frame stalled (no lane could move) and left early");
  return SKW_OK;
}

skw_status skw_batch_check(skw_batch *b) {
  if (!b) return fail(SKW_INVALID_ARG, "batch must not be NULL");
  SKW_HIP(hipDeviceSynchronize(), "synchronize");
  uint64_t stalled = 0;
  SKW_HIP(hipMemcpy(&stalled, b->d_state + 2, sizeof(uint64_t), hipMemcpyDeviceToHost), "state copy");
  if (stalled != 0) return fail(SKW_HIP_ERROR, "a wave of the last batch frame stalled (no lane could move) and left early");
  return SKW_OK;
}

skw_status skw_batch_rays(const skw_batch *b, uint32_t scene, skw_ray *d_rays, void *stream) {
  if (!b || scene >= b->table.size() || !d_rays) return fail(SKW_INVALID_ARG, "bad arguments");
  SKW_HIP(launch_batch_rays(b->d_table + scene, d_rays, b->table[scene].n_rays, reinterpret_cast<hipStream_t>(stream)),
          "ray generation");
  return SKW_OK;
}

skw_status skw_batch_ray_count(const skw_batch *b, uint64_t *n) {
  if (!b || !n) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  *n = b->total_rays;
  return SKW_OK;
}

}  // extern "C"
