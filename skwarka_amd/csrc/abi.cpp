// The C ABI of skwarka_rt (include/skwarka_rt.h): handle ownership, device uploads, argument
// checks, error reporting, and kernel launches. There is no CPU fallback: every query runs on the
// HIP device, and a call without a device fails with SKW_NO_DEVICE.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace skw {

namespace {
thread_local std::string g_err;
}

skw_status fail(skw_status s, const std::string &msg) {
  g_err = msg;
  return s;
}

const char *last_error() { return g_err.c_str(); }

skw_status hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? SKW_OUT_OF_MEMORY : SKW_HIP_ERROR,
              std::string(what) + ": " + hipGetErrorString(e));
}

int device_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atoi(v);
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_pool_mu;
hipMemPool_t g_pools[kMaxDevices] = {};
}  // namespace

hipError_t pool_alloc(void **p, size_t bytes, hipStream_t stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  hipMemPool_t pool = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!g_pools[dev]) {
      hipMemPoolProps props;
      std::memset(&props, 0, sizeof(props));
      props.allocType = hipMemAllocationTypePinned;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      hipMemPool_t created = nullptr;
      e = hipMemPoolCreate(&created, &props);
      if (e != hipSuccess) return e;
      uint64_t keep = UINT64_MAX;
      e = hipMemPoolSetAttribute(created, hipMemPoolAttrReleaseThreshold, &keep);
      if (e != hipSuccess) {
        (void)hipMemPoolDestroy(created);
        return e;
      }
      g_pools[dev] = created;
    }
    pool = g_pools[dev];
  }
  return hipMallocFromPoolAsync(p, bytes, pool, stream);
}

void pool_free(void *p, hipStream_t stream) {
  if (p) (void)hipFreeAsync(p, stream);
}

static int clamp_int(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Lanes a trace wave gathers before it takes new rays / tests leaves (see traverse.hip).
constexpr int kTraceRefillMin = 8;
constexpr int kTraceLeafMin = 16;
constexpr int kTraceClaim = 64;  // rays per claim on the launch counter
constexpr int kTraceWideWavesPerCu = 19;  // five-wave instance, 32-entry stacks (20 for 24)

CounterPool::~CounterPool() {
  for (auto &kv : by_stream) (void)hipFree(kv.second);
}

hipError_t CounterPool::get(void *stream, uint64_t **out) {
  std::lock_guard<std::mutex> lk(mu);
  auto it = by_stream.find(stream);
  if (it != by_stream.end()) {
    *out = it->second;
    return hipSuccess;
  }
  uint64_t *p = nullptr;
  hipError_t e = hipMalloc((void **)&p, sizeof(uint64_t));
  if (e != hipSuccess) return e;
  by_stream[stream] = p;
  *out = p;
  return hipSuccess;
}

skw_status enqueue_trace(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, const uint64_t *n_dev,
                         skw_hit *d_hits, uint8_t *d_occ, hipStream_t st, bool any, uint32_t *d_pops,
                         uint32_t *d_tests, bool fix, uint64_t *d_sums, const TraceStorage *fixed) {
  skw_scene *s = const_cast<skw_scene *>(scene);
  const bool stats = d_pops != nullptr || d_tests != nullptr || d_sums != nullptr;
  uint32_t need = s->host.stack_need;
  uint64_t waves = (n + 63) / 64;
  // A launch with work for that many waves per CU or more, on a tree whose stack fits 32 LDS
  // entries, runs the walk compiled for five waves per SIMD at 20 waves per CU (stacks of up to 24
  // entries, 6,400 bytes of LDS per wave) or 19 (up to 32 entries: 8,448 bytes per wave, of which
  // the CU holds 19); smaller launches (one round of waves) run the instance whose registers do not
  // spill, 16 waves per CU. SKW_TRACE_WAVES_PER_CU above 16 selects the five-wave instance for any
  // launch whose stack fits.
  const bool wide_fits = need <= 32 && !stats;
  const int wide_waves = need <= 24 ? kTraceWideWavesPerCu + 1 : kTraceWideWavesPerCu;
  const bool big = waves >= (uint64_t)s->cus * (uint64_t)wide_waves;
  const int env_w = env_int("SKW_TRACE_WAVES_PER_CU", 0);
  const int waves_per_cu = env_w > 0 ? env_w : (wide_fits && big ? wide_waves : 16);
  const bool wide = wide_fits && waves_per_cu > 16;
  uint64_t cap = (uint64_t)s->cus * (uint64_t)waves_per_cu;
  uint32_t grid = (uint32_t)(waves < cap ? waves : cap);
  if (grid == 0) grid = 1;
  TraceArgs a;
  a.nodes = s->d_nodes;
  a.tris = s->d_tris;
  a.insts = s->d_insts;
  a.inst_range = s->d_inst_range;
  a.root = s->host.root;
  a.rays = d_rays;
  a.hits = d_hits;
  a.occluded = d_occ;
  a.n = n;
  a.n_dev = n_dev;
  a.gstack = nullptr;
  a.pops = d_pops;
  a.tests = d_tests;
  a.sums = d_sums;
  a.refill_min = (uint32_t)clamp_int(env_int("SKW_TRACE_REFILL_MIN", kTraceRefillMin), 1, 64);
  a.leaf_min = (uint32_t)clamp_int(env_int("SKW_TRACE_LEAF_MIN", kTraceLeafMin), 1, 64);
  a.claim = (uint32_t)clamp_int(env_int("SKW_TRACE_CLAIM", kTraceClaim), 64, 4096);
  a.nested = s->host.nested ? 1u : 0u;
  if (fixed) {
    a.counter = fixed->counter;
    if (need > 128) {
      const uint64_t per_wave = (uint64_t)64 * need * sizeof(uint32_t);
      const uint64_t fit = fixed->gstack_bytes / per_wave;
      if (fit == 0) return fail(SKW_UNSUPPORTED, "the tree's traversal stack does not fit the working memory");
      if (grid > fit) grid = (uint32_t)fit;
      a.gstack = fixed->gstack;
    }
  } else {
    SKW_HIP(s->counters.get(st, &a.counter), "counter allocation");
    if (need > 128) {
      SKW_HIP(pool_alloc((void **)&a.gstack, (size_t)grid * 64 * need * sizeof(uint32_t), st), "stack allocation");
    }
  }
  SKW_HIP(launch_set_counter(a.counter, (uint64_t)grid * 64, st), "counter init");
  hipError_t e = any ? launch_any_hit(a, s->has_inst, stats, need, grid, wide, st)
                     : launch_closest_hit(a, s->has_inst, stats, need, grid, wide, st);
  if (!fixed) pool_free(a.gstack, st);
  if (e != hipSuccess) return hip_fail(e, any ? "any_hit launch" : "closest_hit launch");
  if (!any && fix && s->needs_fix) {
    GeomArgs g = s->geom_args();
    g.rays = d_rays;
    g.hits = d_hits;
    g.n = n;
    SKW_HIP(launch_fix_bary(g, st), "barycentric fix-up launch");
  }
  return SKW_OK;
}

GatherArgs gather_args(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                       uint32_t k, double r, uint64_t paths, double rho, int filter) {
  GatherArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nodes = map->d_nodes;
  a.photons = map->d_photons;
  a.psplit = map->d_psplit;
  a.perm = map->d_perm;
  a.cdir = map->d_cdir;
  a.nbox = map->d_nbox;
  a.near = map->d_near;
  a.n_nodes = map->n;
  a.points = d_points;
  a.normals = d_normals;
  a.nq = nq;
  a.k = k;
  a.r = r;
  a.paths = (double)paths;
  a.rho = rho;
  a.sin_t = std::sqrt(1.0 - 0.87 * 0.87);  // PhotonFilter's sin threshold for cos 0.87
  a.filter = filter;
  a.q_stride = k;  // a query's heap slots are contiguous
  a.j_stride = 1;
  a.stack_need = map->max_depth + 1;
  return a;
}

}  // namespace skw

skw_scene::~skw_scene() {
  void *ptrs[] = {d_nodes, d_tris, d_insts, d_inst_range, d_prims, d_meshes, d_mesh_verts, d_mesh_normals, d_mesh_tris};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
}

GeomArgs skw_scene::geom_args() const {
  GeomArgs g;
  g.prims = d_prims;
  g.meshes = d_meshes;
  g.mesh_verts = d_mesh_verts;
  g.mesh_normals = d_mesh_normals;
  g.mesh_tris = d_mesh_tris;
  g.rays = nullptr;
  g.hits = nullptr;
  g.geom = nullptr;
  g.n = 0;
  return g;
}

namespace skw {
skw_status finish_photon_map(skw_photon_map *m) {
  m->n = m->host.nodes.size();
  m->max_depth = m->host.max_depth;
  SKW_HIP(hipGetDevice(&m->device), "hipGetDevice");
  SKW_HIP(upload(&m->d_nodes, m->host.nodes), "upload kd nodes");
  SKW_HIP(upload(&m->d_photons, m->host.photons), "upload photons");
  SKW_HIP(upload(&m->d_psplit, m->host.psplit), "upload parent splits");
  SKW_HIP(upload(&m->d_perm, m->host.perm), "upload permutation");
  SKW_HIP(shared_cdir(&m->d_cdir), "direction table");
  m->cdir_shared = true;
  m->on_device = true;
  return SKW_OK;
}

skw_status photon_map_nbox(skw_photon_map *m, hipStream_t st) {
  if (m->n == 0) return SKW_OK;
  SKW_HIP(hipMalloc((void **)&m->d_nbox, m->n * sizeof(uint64_t)), "normal boxes");
  SKW_HIP(hipMalloc((void **)&m->d_near, m->n * sizeof(DevNearRec)), "nearest-lookup records");
  SKW_HIP(launch_kd_nbox(m->d_nodes, m->d_photons, m->d_cdir, m->n, m->d_nbox, m->d_near, st), "normal boxes");
  SKW_HIP(hipStreamSynchronize(st), "normal boxes");
  return SKW_OK;
}
}  // namespace skw

skw_photon_map::~skw_photon_map() {
  if (pooled) {
    // as hipFree would: no kernel on the device may still read the arrays
    int cur = 0;
    const bool ok = hipGetDevice(&cur) == hipSuccess && hipSetDevice(device) == hipSuccess;
    if (ok) (void)hipDeviceSynchronize();
    for (void *p : {(void *)d_nodes, (void *)d_photons, (void *)d_psplit, (void *)d_perm})
      if (p) skw::pool_free(p, nullptr);
    if (ok) (void)hipSetDevice(cur);
  } else {
    for (void *p : {(void *)d_nodes, (void *)d_photons, (void *)d_psplit, (void *)d_perm})
      if (p) (void)hipFree(p);
  }
  if (d_cdir && !cdir_shared) (void)hipFree(d_cdir);
  if (d_nbox) (void)hipFree(d_nbox);
  if (d_near) (void)hipFree(d_near);
}

extern "C" {

skw_status skw_sort_pairs_u32(const uint32_t *d_keys_in, uint32_t *d_keys_out, const uint32_t *d_vals_in,
                              uint32_t *d_vals_out, uint64_t n, int begin_bit, int end_bit, void *stream) {
  if (n > 0xFFFFFFFFull) return fail(SKW_INVALID_ARG, "at most 2^32 - 1 pairs");
  if (begin_bit < 0 || end_bit > 32 || begin_bit > end_bit) return fail(SKW_INVALID_ARG, "bad key bit range");
  if (n > 0 && (!d_keys_in || !d_keys_out || !d_vals_in || !d_vals_out))
    return fail(SKW_INVALID_ARG, "arrays must not be NULL");
  if (device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  if (n == 0) return SKW_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  size_t bytes = 0;
  SKW_HIP(sort_pairs_u32(nullptr, bytes, nullptr, nullptr, nullptr, nullptr, n, begin_bit, end_bit, st),
          "sort workspace size");
  void *tmp = nullptr;
  SKW_HIP(pool_alloc(&tmp, bytes, st), "sort workspace");
  const hipError_t e = sort_pairs_u32(tmp, bytes, d_keys_in, d_keys_out, d_vals_in, d_vals_out, n, begin_bit,
                                      end_bit, st);
  pool_free(tmp, st);
  if (e != hipSuccess) return hip_fail(e, "radix sort");
  return SKW_OK;
}

const char *skw_last_error(void) { return skw::last_error(); }

int skw_device_count(void) { return device_count(); }

}  // extern "C"

namespace skw {
skw_status scene_tables(skw_scene *s, bool host_only) {
  if (s->tables_done) return SKW_OK;
  std::vector<uint32_t> users(s->host.meshes.size(), 0);
  for (const HostPrim &p : s->host.prims) users[p.mesh]++;
  std::vector<DevPrim> prims(s->host.prims.size());
  for (size_t i = 0; i < prims.size(); ++i) {
    const HostPrim &hp = s->host.prims[i];
    std::memcpy(prims[i].m2w, hp.m2w, sizeof(hp.m2w));
    std::memcpy(prims[i].w2m, hp.w2m, sizeof(hp.w2m));
    prims[i].mesh = hp.mesh;
    // Traversal barycentrics equal the mesh-space ones for identity transforms and for
    // instanced meshes (whose subtrees are walked in mesh space).
    prims[i].identity = (hp.identity || users[hp.mesh] > 1) ? 1u : 0u;
    prims[i].pad[0] = prims[i].pad[1] = 0;
    if (!prims[i].identity) s->needs_fix = true;
  }
  if (host_only) {
    s->tables_done = true;
    return SKW_OK;
  }
  SKW_HIP(hipGetDevice(&s->device), "hipGetDevice");
  SKW_HIP(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device), "hipDeviceGetAttribute");
  std::vector<DevMesh> meshes(s->host.meshes.size());
  std::vector<float> mv, mn;
  std::vector<uint32_t> mt;
  bool any_normals = false;
  for (const HostMesh &m : s->host.meshes) any_normals |= !m.normals.empty();
  for (size_t i = 0; i < meshes.size(); ++i) {
    const HostMesh &m = s->host.meshes[i];
    meshes[i].vert_off = mv.size() / 3;
    meshes[i].tri_off = mt.size() / 3;
    meshes[i].has_normals = m.normals.empty() ? 0u : 1u;
    meshes[i].invert_normals = m.invert_normals ? 1u : 0u;
    mv.insert(mv.end(), m.verts.begin(), m.verts.end());
    if (any_normals) {
      if (m.normals.empty())
        mn.insert(mn.end(), m.verts.size(), 0.0f);
      else
        mn.insert(mn.end(), m.normals.begin(), m.normals.end());
    }
    mt.insert(mt.end(), m.tris.begin(), m.tris.end());
  }
  SKW_HIP(upload(&s->d_prims, prims), "upload primitives");
  SKW_HIP(upload(&s->d_meshes, meshes), "upload meshes");
  SKW_HIP(upload(&s->d_mesh_verts, mv), "upload mesh vertices");
  SKW_HIP(upload(&s->d_mesh_normals, mn), "upload mesh normals");
  SKW_HIP(upload(&s->d_mesh_tris, mt), "upload mesh triangles");
  s->tables_done = true;
  return SKW_OK;
}

// Device tables of a scene whose tree is complete (built on the host or the device, or loaded).
skw_status finish_scene(std::unique_ptr<skw_scene> &sp, bool host_only) {
  skw_scene *s = sp.get();
  s->has_inst = !s->host.insts.empty();
  skw_status st = scene_tables(s, host_only);
  if (st != SKW_OK || host_only) return st;
  // a device-built tree is already in place
  if (!s->d_nodes) SKW_HIP(upload(&s->d_nodes, s->host.nodes), "upload nodes");
  if (!s->d_tris) SKW_HIP(upload(&s->d_tris, s->host.tris), "upload triangles");
  SKW_HIP(upload(&s->d_insts, s->host.insts), "upload instances");
  SKW_HIP(upload(&s->d_inst_range, s->host.inst_range), "upload instance ranges");
  s->on_device = true;
  return SKW_OK;
}
}  // namespace skw

extern "C" {

skw_status skw_scene_create(const skw_scene_desc *desc, uint32_t flags, skw_scene **out) {
  if (!desc || !out) return fail(SKW_INVALID_ARG, "desc and out must not be NULL");
  *out = nullptr;
  bool host_only = (flags & SKW_SCENE_HOST_ONLY) != 0;
  if (!host_only && device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  const bool device_build = (flags & SKW_SCENE_DEVICE_BUILD) != 0;
  if (host_only && device_build) return fail(SKW_INVALID_ARG, "SKW_SCENE_DEVICE_BUILD needs a device");
  std::unique_ptr<skw_scene> s(new skw_scene());
  std::string err;
  bool built = false;
  if (device_build) {
    skw_status bs = build_triangle_tree_device(*desc, s.get(), built);
    if (bs != SKW_OK) return bs;
  }
  if (!built && !build_triangle_tree(*desc, s->host, err)) return fail(SKW_INVALID_ARG, err);
  skw_status st = finish_scene(s, host_only);
  if (st != SKW_OK) return st;
  *out = s.release();
  return SKW_OK;
}

skw_status skw_scene_destroy(skw_scene *scene) {
  delete scene;
  return SKW_OK;
}

skw_status skw_scene_bounds(const skw_scene *scene, double out[6]) {
  if (!scene || !out) return fail(SKW_INVALID_ARG, "scene and out must not be NULL");
  for (int i = 0; i < 6; ++i) out[i] = scene->host.bounds[i];
  return SKW_OK;
}

skw_status skw_scene_info(const skw_scene *scene, uint64_t out[6]) {
  if (!scene || !out) return fail(SKW_INVALID_ARG, "scene and out must not be NULL");
  out[0] = scene->host_tree ? scene->host.nodes.size() : scene->n_nodes;
  out[1] = scene->host_tree ? scene->host.tris.size() : scene->n_tris;
  out[2] = scene->host.max_depth;
  out[3] = scene->host.insts.size();
  out[4] = scene->host.leaves;
  out[5] = scene->host.stack_need;
  return SKW_OK;
}

skw_status skw_scene_export_tree(const skw_scene *scene, float *bbox, uint64_t *meta, uint32_t *tri_prim,
                                 uint32_t *tri_tri, uint64_t *n_nodes) {
  if (!scene) return fail(SKW_INVALID_ARG, "scene must not be NULL");
  if ((tri_prim == nullptr) != (tri_tri == nullptr)) return fail(SKW_INVALID_ARG, "tri_prim and tri_tri go together");
  const TreeLayout *t = nullptr;
  skw_status st = scene_host(scene, &t);
  if (st != SKW_OK) return st;
  uint64_t k = export_tree(*t, bbox, meta, tri_prim, tri_tri);
  if (n_nodes) *n_nodes = k;
  return SKW_OK;
}

static skw_status trace(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits,
                        uint8_t *d_occ, void *stream, bool any, uint32_t *d_pops = nullptr,
                        uint32_t *d_tests = nullptr) {
  if (!scene) return fail(SKW_INVALID_ARG, "scene must not be NULL");
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (n == 0) return SKW_OK;
  if (!d_rays || (!any && !d_hits) || (any && !d_occ)) return fail(SKW_INVALID_ARG, "ray/result pointers must not be NULL");
  return enqueue_trace(scene, d_rays, n, nullptr, d_hits, d_occ, reinterpret_cast<hipStream_t>(stream), any, d_pops,
                       d_tests, true);
}

skw_status skw_closest_hit(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits, void *stream) {
  return trace(scene, d_rays, n, d_hits, nullptr, stream, false);
}

skw_status skw_closest_hit_stats(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits,
                                 uint32_t *d_pops, uint32_t *d_tests, void *stream) {
  return trace(scene, d_rays, n, d_hits, nullptr, stream, false, d_pops, d_tests);
}

skw_status skw_any_hit(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, uint8_t *d_occluded, void *stream) {
  return trace(scene, d_rays, n, nullptr, d_occluded, stream, true);
}

skw_status skw_hit_geometry(const skw_scene *scene, const skw_ray *d_rays, const skw_hit *d_hits, uint64_t n,
                            skw_hit_geom *d_geom, void *stream) {
  if (!scene) return fail(SKW_INVALID_ARG, "scene must not be NULL");
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (n == 0) return SKW_OK;
  if (!d_rays || !d_hits || !d_geom) return fail(SKW_INVALID_ARG, "pointers must not be NULL");
  GeomArgs g;
  g.prims = scene->d_prims;
  g.meshes = scene->d_meshes;
  g.mesh_verts = scene->d_mesh_verts;
  g.mesh_normals = scene->d_mesh_normals;
  g.mesh_tris = scene->d_mesh_tris;
  g.rays = d_rays;
  g.hits = const_cast<skw_hit *>(d_hits);
  g.geom = d_geom;
  g.n = n;
  SKW_HIP(launch_hit_geometry(g, reinterpret_cast<hipStream_t>(stream)), "hit_geometry launch");
  return SKW_OK;
}

skw_status skw_photon_map_create(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                                 const uint16_t *normal, uint32_t flags, skw_photon_map **out) {
  if (!out) return fail(SKW_INVALID_ARG, "out must not be NULL");
  *out = nullptr;
  if (n > 0 && !pos) return fail(SKW_INVALID_ARG, "pos must not be NULL");
  if (n >= (uint64_t)kNoChild) return fail(SKW_INVALID_ARG, "too many photons for the 29-bit child index");
  bool host_only = (flags & SKW_SCENE_HOST_ONLY) != 0;
  if (!host_only && device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  std::unique_ptr<skw_photon_map> m(new skw_photon_map());
  build_photon_tree(n, pos, weight, dir, normal, m->host);
  if (!host_only) {
    skw_status st = finish_photon_map(m.get());
    if (st != SKW_OK) return st;
  } else {
    m->n = n;
    m->max_depth = m->host.max_depth;
  }
  *out = m.release();
  return SKW_OK;
}

skw_status skw_photon_map_destroy(skw_photon_map *map) {
  delete map;
  return SKW_OK;
}

skw_status skw_photon_map_info(const skw_photon_map *map, uint64_t *n, uint32_t *max_depth) {
  if (!map) return fail(SKW_INVALID_ARG, "map must not be NULL");
  if (n) *n = map->n;
  if (max_depth) *max_depth = map->max_depth;
  return SKW_OK;
}

skw_status skw_photon_map_export(const skw_photon_map *map, uint32_t *perm, float *split, uint32_t *word) {
  if (!map) return fail(SKW_INVALID_ARG, "map must not be NULL");
  const KdLayout &h = photon_map_host(map);
  if (h.nodes.size() != map->n) return fail(SKW_HIP_ERROR, "photon tree download failed");
  for (size_t i = 0; i < h.nodes.size(); ++i) {
    if (perm) perm[i] = h.perm[i];
    uint32_t w = h.nodes[i].word;
    if (word) word[i] = w;
    if (split) {
      uint32_t axis = w & 3u;
      split[i] = axis == 3u ? 0.0f : h.nodes[i].p[axis];
    }
  }
  return SKW_OK;
}

// Estimator choice of a k-nearest gather (GatherArgs::estimate): 0 disc, 1 cone, 2 caustic.
struct Estimate {
  int kind = 0;
  double kc = 0.0;
  const double *wo = nullptr;
  double *caustic = nullptr;
  int ids_tree = 0;  // ids hold caller indices (0), tree positions (1) or direction codes (2)
  const uint32_t *order = nullptr;  // visiting order of the queries (device), optional
  uint32_t *stats = nullptr;        // per query (visits, photons inside the radius, set size)
};

static skw_status gather(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                         uint32_t k, double r, uint64_t paths, double rho, int filter, int mode, float *e_ext,
                         float *e_int, float *radiance, uint32_t *ids, double *d2, uint32_t *count,
                         uint32_t *nearest, void *stream, const Estimate &est = Estimate()) {
  if (!map) return fail(SKW_INVALID_ARG, "map must not be NULL");
  if (!map->on_device) return fail(SKW_UNSUPPORTED, "photon map was created host-only");
  if (!(r >= 0.0)) return fail(SKW_INVALID_ARG, "radius must be >= 0");
  if (filter != 0 && filter != 1) return fail(SKW_INVALID_ARG, "filter must be 0 or 1");
  if (nq == 0) return SKW_OK;
  if (!d_points) return fail(SKW_INVALID_ARG, "points must not be NULL");
  if (filter == 1 && !d_normals) return fail(SKW_INVALID_ARG, "PhotonFilter needs normals");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  GatherArgs a = gather_args(map, d_points, d_normals, nq, k, r, paths, rho, filter);
  a.ids = ids;
  a.d2 = d2;
  a.count = count;
  a.e_ext = e_ext;
  a.e_int = e_int;
  a.radiance = radiance;
  a.nearest = nearest;
  a.estimate = est.kind;
  a.kc = est.kc;
  a.wo = est.wo;
  a.caustic = est.caustic;
  a.ids_tree = est.ids_tree;
  a.order = est.order;
  a.stats = est.stats;
  // A disc estimate whose set is not returned writes the set's slots only if the set fills.
  a.defer = (mode == 0 && est.kind == 0 && e_ext && !ids && !d2) ? 1 : 0;
  if (mode == 0 && k > 0) {
    size_t elems = (size_t)nq * k;
    SKW_HIP(pool_alloc((void **)&a.heap_pos, elems * sizeof(uint32_t), st), "heap allocation");
    hipError_t e = pool_alloc((void **)&a.heap_d2, elems * sizeof(double), st);
    if (e != hipSuccess) {
      pool_free(a.heap_pos, st);
      return hip_fail(e, "heap allocation");
    }
  }
  if (a.stack_need > 64) {
    hipError_t e = pool_alloc((void **)&a.gstack, (size_t)nq * a.stack_need * sizeof(uint32_t), st);
    if (e != hipSuccess) {
      pool_free(a.heap_pos, st);
      pool_free(a.heap_d2, st);
      return hip_fail(e, "stack allocation");
    }
  }
  hipError_t e = launch_gather(a, mode, st);
  pool_free(a.heap_pos, st);
  pool_free(a.heap_d2, st);
  pool_free(a.gstack, st);
  if (e != hipSuccess) return hip_fail(e, "gather launch");
  return SKW_OK;
}

}  // extern "C"

namespace skw {
skw_status photon_knn_codes(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                            uint32_t k, double r, uint32_t *d_codes, uint32_t *d_count, hipStream_t stream) {
  Estimate est;
  est.ids_tree = 2;
  return gather(map, d_points, d_normals, nq, k, r, 1, 0.0, 1, 0, nullptr, nullptr, nullptr, d_codes, nullptr, d_count,
                nullptr, stream, est);
}

skw_status photon_nearest_ordered(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                  uint64_t nq, double r, const uint32_t *d_order, uint32_t *d_id, hipStream_t stream) {
  Estimate est;
  est.order = d_order;
  return gather(map, d_points, d_normals, nq, 0, r, 1, 0.0, 1, 1, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, d_id, stream, est);
}
}  // namespace skw

extern "C" {

skw_status skw_photon_irradiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                 uint64_t nq, uint32_t k, double r, uint64_t paths, double rho, float *d_e_ext,
                                 float *d_e_int, float *d_radiance, uint32_t *d_ids, double *d_d2,
                                 uint32_t *d_count, void *stream) {
  if (nq > 0 && (!d_e_ext || !d_e_int)) return fail(SKW_INVALID_ARG, "e_ext and e_int must not be NULL");
  if (nq > 0 && !d_normals) return fail(SKW_INVALID_ARG, "normals must not be NULL");
  return gather(map, d_points, d_normals, nq, k, r, paths, rho, 1, 0, d_e_ext, d_e_int, d_radiance, d_ids, d_d2,
                d_count, nullptr, stream);
}

skw_status skw_photon_irradiance_stats(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                       uint64_t nq, uint32_t k, double r, uint64_t paths, double rho, float *d_e_ext,
                                       float *d_e_int, float *d_radiance, uint32_t *d_stats, void *stream) {
  if (nq > 0 && (!d_e_ext || !d_e_int || !d_stats)) return fail(SKW_INVALID_ARG, "e_ext, e_int and stats must not be NULL");
  if (nq > 0 && !d_normals) return fail(SKW_INVALID_ARG, "normals must not be NULL");
  Estimate est;
  est.stats = d_stats;
  return gather(map, d_points, d_normals, nq, k, r, paths, rho, 1, 0, d_e_ext, d_e_int, d_radiance, nullptr, nullptr,
                nullptr, nullptr, stream, est);
}

skw_status skw_photon_cone_irradiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                      uint64_t nq, uint32_t k, double r, uint64_t paths, double kc, float *d_e_ext,
                                      void *stream) {
  if (nq > 0 && (!d_e_ext || !d_normals)) return fail(SKW_INVALID_ARG, "normals and e_ext must not be NULL");
  if (!(kc >= 1.0)) return fail(SKW_INVALID_ARG, "cone constant kc must be >= 1");
  if (k == 0) return fail(SKW_INVALID_ARG, "k must be > 0");
  Estimate est;
  est.kind = 1;
  est.kc = kc;
  return gather(map, d_points, d_normals, nq, k, r, paths, 0.0, 1, 0, d_e_ext, nullptr, nullptr, nullptr, nullptr,
                nullptr, nullptr, stream, est);
}

skw_status skw_photon_caustic_radiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                       const double *d_wo, uint64_t nq, uint32_t k, double r, uint64_t paths,
                                       double rho, double *d_radiance, void *stream) {
  if (nq > 0 && (!d_radiance || !d_normals || !d_wo))
    return fail(SKW_INVALID_ARG, "normals, wo and radiance must not be NULL");
  if (k == 0) return fail(SKW_INVALID_ARG, "k must be > 0");
  Estimate est;
  est.kind = 2;
  est.wo = d_wo;
  est.caustic = d_radiance;
  return gather(map, d_points, d_normals, nq, k, r, paths, rho, 1, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, nullptr, stream, est);
}

skw_status skw_photon_knn(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                          uint32_t k, double r, int filter, uint32_t *d_ids, double *d_d2, uint32_t *d_count,
                          void *stream) {
  if (nq > 0 && (!d_ids || !d_count)) return fail(SKW_INVALID_ARG, "ids and count must not be NULL");
  return gather(map, d_points, d_normals, nq, k, r, 1, 0.0, filter, 0, nullptr, nullptr, nullptr, d_ids, d_d2,
                d_count, nullptr, stream);
}

skw_status skw_photon_nearest(const skw_photon_map *map, const double *d_points, const double *d_normals,
                              uint64_t nq, double r, int filter, uint32_t *d_id, void *stream) {
  if (nq > 0 && !d_id) return fail(SKW_INVALID_ARG, "id must not be NULL");
  return gather(map, d_points, d_normals, nq, 0, r, 1, 0.0, filter, 1, nullptr, nullptr, nullptr, nullptr, nullptr,
                nullptr, d_id, stream);
}

}  // extern "C"
