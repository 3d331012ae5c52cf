// C ABI of photon shooting (skw_photon_shoot, skw_photon_set_*): the host loop over rounds and
// bounces around the kernels of shoot.hip.
//
// PhotonLTEIntegrator::ShootPhotons (Raytracer/LTEIntegrators/PhotonLTEIntegrator/
// PhotonLTEIntegrator.cpp:247-288) with the input filter's stop rules and the merging filter's
// per-map bookkeeping (PhotonShootingPipeline.cpp:48-92, 254-310), one round of paths at a time;
// the lights' power CDF of _GetLightsPowerCDF (PhotonLTEIntegrator.cpp:80-124).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"

using namespace skw;

struct skw_dl_scene;

namespace skw {
void dl_tables(const skw_dl_scene *dl, const DevMat **mats, const double **rad);
void dl_lights(const skw_dl_scene *dl, const std::vector<DevLight> **host, const DevLight **d_lights,
               const DevLightTri **d_ltris);
}  // namespace skw

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDblEps = 2.220446049250313080847e-16;
const double kLumW[3] = {0x1.b38293da589a6p-5, 0x1.2b993cc221566p-1, 0x1.725d3400723f6p-2};

// A growable device array of photon records (SoA).
struct PhotonArrays {
  float *pos = nullptr, *w = nullptr;
  uint32_t *codes = nullptr;
  uint64_t n = 0, cap = 0;
  void release() {
    if (pos) (void)hipFree(pos);
    if (w) (void)hipFree(w);
    if (codes) (void)hipFree(codes);
    pos = w = nullptr;
    codes = nullptr;
    n = cap = 0;
  }
  // Room for `need` records, keeping the first n.
  hipError_t reserve(uint64_t need, hipStream_t st, uint32_t **key = nullptr, uint32_t **key_old = nullptr) {
    if (need <= cap) return hipSuccess;
    uint64_t c = std::max<uint64_t>(need, cap * 2);
    float *np = nullptr, *nw = nullptr;
    uint32_t *nc = nullptr;
    hipError_t e = hipMalloc((void **)&np, c * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&nw, c * 3 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&nc, c * sizeof(uint32_t));
    if (e == hipSuccess && n > 0) {
      e = hipMemcpyAsync(np, pos, n * 3 * sizeof(float), hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(nw, w, n * 3 * sizeof(float), hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(nc, codes, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      if (np) (void)hipFree(np);
      if (nw) (void)hipFree(nw);
      if (nc) (void)hipFree(nc);
      return e;
    }
    const uint64_t keep = n;
    release();
    pos = np;
    w = nw;
    codes = nc;
    n = keep;
    cap = c;
    return hipSuccess;
  }
};

}  // namespace

struct skw_photon_set {
  PhotonArrays m[3];
  uint64_t paths[3] = {0, 0, 0};
  uint64_t total_paths = 0;
  ~skw_photon_set() {
    for (auto &a : m) a.release();
  }
};

namespace {

// The round's photon buffers of one map: records plus their path keys, growable.
struct RoundPhotons {
  PhotonArrays rec;
  uint32_t *key = nullptr;
  ~RoundPhotons() {
    rec.release();
    if (key) (void)hipFree(key);
  }
  hipError_t reserve(uint64_t need, hipStream_t st) {
    if (need <= rec.cap) return hipSuccess;
    const uint64_t old_n = rec.n;
    hipError_t e = rec.reserve(need, st);
    if (e != hipSuccess) return e;
    uint32_t *nk = nullptr;
    e = hipMalloc((void **)&nk, rec.cap * sizeof(uint32_t));
    if (e == hipSuccess && old_n > 0 && key)
      e = hipMemcpyAsync(nk, key, old_n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      if (nk) (void)hipFree(nk);
      return e;
    }
    if (key) (void)hipFree(key);
    key = nk;
    return hipSuccess;
  }
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(uint64_t n) { return hipMalloc((void **)&p, std::max<uint64_t>(n, 1) * sizeof(T)); }
};

}  // namespace

extern "C" {

skw_status skw_photon_shoot(const skw_scene *scene, const skw_dl_scene *dl, const skw_shoot_params *prm,
                            skw_photon_set **out, void *stream) {
  if (!scene || !dl || !prm || !out) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  *out = nullptr;
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (prm->paths_per_round == 0) return fail(SKW_INVALID_ARG, "paths_per_round must be > 0");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::unique_ptr<skw_photon_set> set(new skw_photon_set());
  const std::vector<DevLight> *lights = nullptr;
  const DevLight *d_lights = nullptr;
  const DevLightTri *d_ltris = nullptr;
  dl_lights(dl, &lights, &d_lights, &d_ltris);
  const DevMat *mats = nullptr;
  const double *rad = nullptr;
  dl_tables(dl, &mats, &rad);
  const uint32_t nl = (uint32_t)lights->size();
  if (nl == 0 || prm->max_paths == 0) {
    *out = set.release();
    return SKW_OK;
  }
  // _GetLightsPowerCDF: luminance of (pi * area) * radiance, accumulated, divided by the total.
  std::vector<double> cdf(nl, 0.0);
  for (uint32_t l = 0; l < nl; ++l) {
    double y = 0.0;
    for (int c = 0; c < 3; ++c) y += kLumW[c] * ((kPi * (*lights)[l].area) * (*lights)[l].radiance[c]);
    cdf[l] = y;
  }
  for (uint32_t i = 1; i < nl; ++i) cdf[i] += cdf[i - 1];
  const double total = cdf[nl - 1];
  if (total > kDblEps)
    for (uint32_t i = 0; i < nl; ++i) cdf[i] /= total;
  else
    cdf[nl - 1] = 1.0;

  const uint64_t R = std::min<uint64_t>(prm->paths_per_round, prm->max_paths);
  DevBuf<double> d_cdf;
  DevBuf<skw_ray> rays[2];
  DevBuf<ShootPath> state[2];
  DevBuf<skw_hit> hits;
  DevBuf<uint64_t> counts;
  DevBuf<uint32_t> iota, order;
  DevBuf<uint32_t> keys_sorted;
  SKW_HIP(d_cdf.alloc(nl), "shooting buffers");
  for (int i = 0; i < 2; ++i) {
    SKW_HIP(rays[i].alloc(R), "shooting buffers");
    SKW_HIP(state[i].alloc(R), "shooting buffers");
  }
  SKW_HIP(hits.alloc(R), "shooting buffers");
  SKW_HIP(counts.alloc(5), "shooting buffers");
  SKW_HIP(hipMemcpyAsync(d_cdf.p, cdf.data(), nl * sizeof(double), hipMemcpyHostToDevice, st), "upload light CDF");
  RoundPhotons round[3];
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  uint64_t sort_cap = 0;
  int key_bits = 1;
  while ((1ull << key_bits) < R) ++key_bits;

  ShootArgs a;
  std::memset(&a, 0, sizeof(a));
  a.geom = scene->geom_args();
  a.mats = mats;
  a.lights = d_lights;
  a.ltris = d_ltris;
  a.light_cdf = d_cdf.p;
  a.n_lights = nl;
  a.seed = prm->seed;
  for (int i = 0; i < 2; ++i) {
    a.rays[i] = rays[i].p;
    a.state[i] = state[i].p;
  }
  a.hits = hits.p;
  a.counts = counts.p;

  skw_status rs = SKW_OK;
  hipError_t e = hipSuccess;
  while (rs == SKW_OK && e == hipSuccess) {
    bool done[3];
    for (int j = 0; j < 3; ++j) done[j] = set->m[j].n >= prm->max_photons[j];
    if (set->total_paths >= prm->max_paths || (done[0] && done[1] && done[2])) break;
    bool stuck = set->total_paths > 10000000ull;
    for (int j = 0; j < 3; ++j) stuck = stuck && (done[j] || set->m[j].n == 0);
    if (stuck) break;
    const uint64_t first = set->total_paths;
    const uint64_t n = std::min<uint64_t>(R, prm->max_paths - first);
    set->total_paths += n;
    a.first = first;
    a.n = n;
    for (int j = 0; j < 3; ++j) {
      a.done[j] = done[j] ? 1u : 0u;
      round[j].rec.n = 0;
    }
    // direct photons: at most one per path
    if (!done[1] && (e = round[1].reserve(n, st)) != hipSuccess) break;
    if ((e = hipMemsetAsync(counts.p, 0, 5 * sizeof(uint64_t), st)) != hipSuccess) break;
    a.cur = 0;
    for (int j = 0; j < 3; ++j) {
      a.ph_pos[j] = round[j].rec.pos;
      a.ph_w[j] = round[j].rec.w;
      a.ph_codes[j] = round[j].rec.codes;
      a.ph_key[j] = round[j].key;
    }
    if ((e = launch_shoot_start(a, st)) != hipSuccess) break;
    uint64_t c[5];
    while (true) {
      if ((e = hipMemcpyAsync(c, counts.p, sizeof(c), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
      const uint64_t n_active = c[a.cur];
      if (n_active == 0) break;
      // every active path may deposit one photon: room in the maps that take photons
      for (int j = 0; j < 3 && e == hipSuccess; ++j) {
        if (done[j] || j == 1) continue;
        round[j].rec.n = c[2 + j];  // records a reallocation keeps
        e = round[j].reserve(c[2 + j] + n_active, st);
      }
      if (e != hipSuccess) break;
      for (int j = 0; j < 3; ++j) {
        a.ph_pos[j] = round[j].rec.pos;
        a.ph_w[j] = round[j].rec.w;
        a.ph_codes[j] = round[j].rec.codes;
        a.ph_key[j] = round[j].key;
      }
      rs = enqueue_trace(scene, rays[a.cur].p, n_active, nullptr, hits.p, nullptr, st, false, nullptr, nullptr, false);
      if (rs != SKW_OK) break;
      if ((e = hipMemsetAsync(counts.p + (a.cur ^ 1u), 0, sizeof(uint64_t), st)) != hipSuccess) break;
      if ((e = launch_shoot_bounce(a, n_active, st)) != hipSuccess) break;
      a.cur ^= 1u;
    }
    if (rs != SKW_OK || e != hipSuccess) break;
    // the round's photons in (path, intersection) order, appended to their maps
    for (int j = 0; j < 3 && e == hipSuccess; ++j) {
      if (done[j]) continue;
      set->paths[j] += n;
      const uint64_t k = c[2 + j];
      if (k == 0) continue;
      if (k > sort_cap) {
        if (sort_tmp) (void)hipFree(sort_tmp);
        sort_tmp = nullptr;
        if (iota.p) (void)hipFree(iota.p);
        if (order.p) (void)hipFree(order.p);
        if (keys_sorted.p) (void)hipFree(keys_sorted.p);
        iota.p = order.p = keys_sorted.p = nullptr;
        sort_cap = std::max<uint64_t>(k, 2 * sort_cap);
        sort_bytes = 0;
        if ((e = sort_pairs_u32(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, sort_cap, 0, key_bits, st)) !=
            hipSuccess)
          break;
        if ((e = hipMalloc(&sort_tmp, std::max<size_t>(sort_bytes, 1))) != hipSuccess) break;
        if ((e = iota.alloc(sort_cap)) != hipSuccess) break;
        if ((e = order.alloc(sort_cap)) != hipSuccess) break;
        if ((e = keys_sorted.alloc(sort_cap)) != hipSuccess) break;
      }
      if ((e = launch_iota_u32(iota.p, k, st)) != hipSuccess) break;
      size_t bytes = sort_bytes;
      if ((e = sort_pairs_u32(sort_tmp, bytes, round[j].key, keys_sorted.p, iota.p, order.p, k, 0, key_bits, st)) !=
          hipSuccess)
        break;
      PhotonArrays &dst = set->m[j];
      if ((e = dst.reserve(dst.n + k, st)) != hipSuccess) break;
      e = launch_photon_gather_records(round[j].rec.pos, round[j].rec.w, round[j].rec.codes, order.p, k,
                                       dst.pos + 3 * dst.n, dst.w + 3 * dst.n, dst.codes + dst.n, st);
      dst.n += k;
    }
  }
  if (e == hipSuccess && rs == SKW_OK) e = hipStreamSynchronize(st);
  if (sort_tmp) (void)hipFree(sort_tmp);
  if (rs != SKW_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "photon shooting");
  *out = set.release();
  return SKW_OK;
}

skw_status skw_photon_set_info(const skw_photon_set *set, uint64_t counts[3], uint64_t paths[3],
                               uint64_t *total_paths) {
  if (!set) return fail(SKW_INVALID_ARG, "set must not be NULL");
  for (int j = 0; j < 3; ++j) {
    if (counts) counts[j] = set->m[j].n;
    if (paths) paths[j] = set->paths[j];
  }
  if (total_paths) *total_paths = set->total_paths;
  return SKW_OK;
}

skw_status skw_photon_set_export(const skw_photon_set *set, uint32_t which, float *pos, float *weight, uint16_t *dir,
                                 uint16_t *normal) {
  if (!set) return fail(SKW_INVALID_ARG, "set must not be NULL");
  if (which > 2) return fail(SKW_INVALID_ARG, "which must be 0, 1 or 2");
  const PhotonArrays &m = set->m[which];
  if (m.n == 0) return SKW_OK;
  if (pos) SKW_HIP(hipMemcpy(pos, m.pos, m.n * 3 * sizeof(float), hipMemcpyDeviceToHost), "photon export");
  if (weight) SKW_HIP(hipMemcpy(weight, m.w, m.n * 3 * sizeof(float), hipMemcpyDeviceToHost), "photon export");
  if (dir || normal) {
    std::vector<uint32_t> codes(m.n);
    SKW_HIP(hipMemcpy(codes.data(), m.codes, m.n * sizeof(uint32_t), hipMemcpyDeviceToHost), "photon export");
    for (uint64_t i = 0; i < m.n; ++i) {
      if (dir) dir[i] = (uint16_t)(codes[i] & 0xFFFFu);
      if (normal) normal[i] = (uint16_t)(codes[i] >> 16);
    }
  }
  return SKW_OK;
}

skw_status skw_photon_set_map(const skw_photon_set *set, uint32_t which, skw_photon_map **out, void *stream) {
  if (!set || !out) return fail(SKW_INVALID_ARG, "set and out must not be NULL");
  if (which > 2) return fail(SKW_INVALID_ARG, "which must be 0, 1 or 2");
  const PhotonArrays &m = set->m[which];
  return photon_map_build_device(m.n, m.pos, m.w, m.codes, nullptr, nullptr, out, reinterpret_cast<hipStream_t>(stream));
}

skw_status skw_photon_set_destroy(skw_photon_set *set) {
  delete set;
  return SKW_OK;
}

}  // extern "C"
