// Internal declarations shared by the C ABI translation units: handle structs, error helpers and
// the trace enqueue used by every caller of the traversal kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/skwarka_rt.h"
#include "host/builders.hpp"
#include "kernels/kernels.hpp"

namespace skw {

skw_status fail(skw_status s, const std::string &msg);
const char *last_error();
skw_status hip_fail(hipError_t e, const char *what);
int device_count();
int env_int(const char *name, int dflt);

#define SKW_HIP(call, what)                            \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return skw::hip_fail(e_, what); \
  } while (0)

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src) {
  *dst = nullptr;
  if (src.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **)dst, src.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Stream-ordered device memory from a pool the library owns, one per device, that keeps freed
// blocks for reuse (release threshold: never). A frame's working set (about 1.3 GB for a
// 1024x1024 direct-lighting pass) is then mapped once per process instead of once per call.
hipError_t pool_alloc(void **p, size_t bytes, hipStream_t stream);
void pool_free(void *p, hipStream_t stream);

// Launch-wide ray counters, one per stream so that calls on different streams never share one.
struct CounterPool {
  std::mutex mu;
  std::map<void *, uint64_t *> by_stream;
  ~CounterPool();
  hipError_t get(void *stream, uint64_t **out);
};

}  // namespace skw

struct skw_scene {
  skw::TreeLayout host;     // nodes / tris pending (host_tree false) for device-built trees
  bool host_tree = true;    // host.nodes / host.tris hold the tree (device builds: on first use)
  uint64_t n_nodes = 0, n_tris = 0;  // counts while the host copy is pending
  std::mutex host_mu;
  bool tables_done = false; // primitive / mesh tables computed (and uploaded unless host-only)
  bool on_device = false;
  bool has_inst = false;
  bool needs_fix = false;  // some primitive's traversal barycentrics are not mesh-space values
  int device = 0;
  int cus = 256;
  skw::DevNode *d_nodes = nullptr;
  skw::DevTri *d_tris = nullptr;
  skw::DevInst *d_insts = nullptr;
  uint32_t *d_inst_range = nullptr;
  skw::DevPrim *d_prims = nullptr;
  skw::DevMesh *d_meshes = nullptr;
  float *d_mesh_verts = nullptr;
  float *d_mesh_normals = nullptr;
  uint32_t *d_mesh_tris = nullptr;
  skw::CounterPool counters;
  ~skw_scene();
  skw::GeomArgs geom_args() const;
};

struct skw_photon_map {
  skw::KdLayout host;       // complete unless the tree was built on the device (see photon_map_host)
  uint64_t n = 0;           // photons
  uint32_t max_depth = 0;   // depth of the deepest node (the root is 0)
  bool host_ready = true;   // `host` holds the tree (device-built maps fill it on first use)
  std::mutex host_mu;
  bool on_device = false;
  int device = 0;
  skw::DevKdNode *d_nodes = nullptr;
  skw::DevPhoton *d_photons = nullptr;
  float *d_psplit = nullptr;
  uint32_t *d_perm = nullptr;
  double *d_cdir = nullptr;
  uint64_t *d_nbox = nullptr;  // per node: its subtree's normal box (irradiance maps; launch_kd_nbox)
  skw::DevNearRec *d_near = nullptr;  // per node: the nearest-lookup record (irradiance maps)
  bool pooled = false;      // arrays from the device pool (device builds)
  bool cdir_shared = false; // d_cdir is the device's shared direction table
  ~skw_photon_map();
};

struct skw_irradiance_map {
  skw_photon_map *map = nullptr;  // tree over the irradiance photons (weight: external irradiance)
  std::vector<float> pos, e_ext, e_int;
  std::vector<uint16_t> normal;
  double max_dist = __builtin_inf();
  float *d_e_ext = nullptr, *d_e_int = nullptr;  // creation order
  ~skw_irradiance_map() {
    if (map) skw_photon_map_destroy(map);
    if (d_e_ext) (void)hipFree(d_e_ext);
    if (d_e_int) (void)hipFree(d_e_int);
  }
};

namespace skw {
// Enqueues closest-hit (any = false) or occlusion (any = true) for n rays; when n_dev is set the
// count is read on the device and n only bounds the launch size. fix: rewrite b1/b2 of hits on
// transformed primitives in mesh space. d_sums (optional, device): [0] += node boxes tested,
// [1] += triangles tested over all rays (selects the counting variant of the walk).
// fixed (optional): caller-owned working memory instead of the scene's counter and the pool: the
// launch counter, and for trees deeper than the LDS stacks a global stack area of gstack_bytes,
// to which the launch's grid is then limited (the kernels are persistent, so any grid completes).
struct TraceStorage {
  uint64_t *counter;
  uint32_t *gstack;
  uint64_t gstack_bytes;
};
skw_status enqueue_trace(const skw_scene *s, const skw_ray *d_rays, uint64_t n, const uint64_t *n_dev, skw_hit *d_hits,
                         uint8_t *d_occ, hipStream_t st, bool any, uint32_t *d_pops, uint32_t *d_tests, bool fix,
                         uint64_t *d_sums = nullptr, const TraceStorage *fixed = nullptr);

// Gather arguments for one photon map and query set, outputs unset and no heap or stack storage
// attached (heap: nq * k slots, slot-major; stack: only when the tree is deeper than 64).
GatherArgs gather_args(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                       uint32_t k, double r, uint64_t paths, double rho, int filter);

// A scene without instanced meshes with its tree built on the current device, node for node the
// host builder's tree, straight into the scene's device arrays (built set; an empty scene is left
// to the host builder); SKW_UNSUPPORTED for instanced scenes or non-finite triangle boxes.
skw_status build_triangle_tree_device(const skw_scene_desc &desc, skw_scene *s, bool &built);
// The scene's host tree (downloaded on first use for device-built trees).
skw_status scene_host(const skw_scene *s, const TreeLayout **out);
// The primitive and mesh tables of a scene whose meshes and primitives are read (uploaded unless
// host_only); once per scene.
skw_status scene_tables(skw_scene *s, bool host_only);

// The per-node normal boxes of a map whose device arrays are complete (synchronises `st`).
skw_status photon_map_nbox(skw_photon_map *m, hipStream_t st);

// Device tables of a scene / photon map whose host layout is complete (built or loaded).
skw_status finish_scene(std::unique_ptr<skw_scene> &s, bool host_only);
skw_status finish_photon_map(skw_photon_map *m);
skw_status upload_irradiance(skw_irradiance_map *m);  // device copies of the irradiance arrays

// The map's tree in host memory, downloaded on first use for device-built maps.
const KdLayout &photon_map_host(const skw_photon_map *map);

// k nearest photons with PhotonFilter within r, each given by its direction code (the low 16 bits
// of its packed ids; row-major, k per query, in the bounded heap's final order).
skw_status photon_knn_codes(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                            uint32_t k, double r, uint32_t *d_codes, uint32_t *d_count, hipStream_t stream);
// A photon map whose kd-tree is built on the device from device arrays (positions, weights, and
// either codes = dir | normal << 16 or the two id arrays; weights / ids may be NULL).
// The direction table of CompressedDirection on the current device, uploaded once and shared.
hipError_t shared_cdir(double **out);
// A second stream of the current device (default priority, non-blocking), kept for the life of the
// process, on which a wavefront runs one of two independent stages beside the caller's stream.
hipError_t side_stream(hipStream_t *out);
skw_status photon_map_build_device(uint64_t n, const float *d_pos, const float *d_weight, const uint32_t *d_codes,
                                   const uint16_t *d_dir, const uint16_t *d_normal, skw_photon_map **out,
                                   hipStream_t st);

// Nearest photon with PhotonFilter within r of each query; thread i serves query d_order[i].
skw_status photon_nearest_ordered(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                  uint64_t nq, double r, const uint32_t *d_order, uint32_t *d_id, hipStream_t stream);
}  // namespace skw
