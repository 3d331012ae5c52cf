// C ABI of the photon final gather (skw_irradiance_map_*, skw_dl_final_gather).
//
// The irradiance photon map follows PhotonLTEIntegrator::_ConstructIrradiancePhotonMap
// (Raytracer/LTEIntegrators/PhotonLTEIntegrator.cpp:171-245): its photons are every 10th photon of
// the indirect map's GetAllPoints (tree order), their irradiance the float sum of three k = 100
// disc estimates (caustic, direct, indirect; IrradiancePhotonProcess, PhotonInternalTypes.h:
// 269-288) run by the gather kernel, and the lookup distance the 0.1 % percentile of the
// nearest-irradiance-photon distances of every 11th indirect photon (:226-244). The final gather
// is a wavefront over each pixel batch of a sample pass:
//   camera rays -> closest hit -> shading-point records -> 32 nearest indirect photons (tree
//   positions) -> BSDF and cone samples appending gather rays -> closest hit of the gather rays ->
//   hit scattering and lookup points -> nearest irradiance photon -> per-sample ordered sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <vector>

#include "internal.hpp"

using namespace skw;

struct skw_dl_scene;  // defined in abi_dl.cpp; only its device tables are read, through dl_tables

namespace skw {
// Device tables of a direct-lighting scene (abi_dl.cpp).
void dl_tables(const skw_dl_scene *dl, const DevMat **mats, const double **rad);
// Entries of a direct-lighting scene's material table (one per primitive; abi_dl.cpp).
uint32_t dl_material_count(const skw_dl_scene *dl);
}  // namespace skw


namespace {

constexpr double kInvPi = 0.31830988618379067154;
constexpr uint32_t kFgK = 32;          // indirect photons per final-gather lookup (PhotonLTEIntegrator.cpp:459)
constexpr uint32_t kMaxGather = 676;   // 26^2: the sample stream has 4096 dims per sample, 6 per gather sample
// Gather rays held per launch, at most (SKW_FG_GATHER_RAYS_LOG2 overrides); fewer when the
// launch's buffers would exceed a quarter of the device's free memory. Ray indices within a launch
// are int32. 2^27 rays hold about 21 GB; at 2^28 (about 41 GB) a c7 frame took 366 ms in the
// first processes of a GPU session but 750 ms in later ones (the device's free memory unchanged),
// while 2^27 and 2^26 stayed steady (375 ms, 388 ms) over repeated processes.
constexpr int kMaxGatherRaysLog2 = 27;
constexpr uint32_t kMaxPassesPerLaunchLog2 = 4;
// Bits per axis of the lookup points' Morton cells: 256^3 cells over the scene box, a 24-bit key
// (three 8-bit sort passes). Finer cells keep a wave's 64 lookups closer together than 64^3 did
// (c7 374.9 -> 361.7 ms); 512^3 gains nothing more.
constexpr int kMortonBits = 8;

double inf() { return std::numeric_limits<double>::infinity(); }

// Positions (as doubles) and decoded normals of the photons at tree positions 0, step, 2 step, ...
void strided_queries(const KdLayout &h, size_t step, std::vector<double> &pts, std::vector<double> &nrm) {
  const double *cdir = cdir_table();
  const size_t n = h.nodes.size(), nq = (n + step - 1) / step;
  pts.assign(3 * nq, 0.0);
  nrm.assign(3 * nq, 0.0);
  for (size_t i = 0; i < nq; ++i) {
    const size_t t = i * step;
    const double *dv = cdir + 3 * (size_t)(h.photons[t].codes >> 16);
    for (int c = 0; c < 3; ++c) {
      pts[3 * i + c] = (double)h.nodes[t].p[c];
      nrm[3 * i + c] = dv[c];
    }
  }
}

// One stream-ordered device block cut into 256-byte aligned pieces.
struct Block {
  size_t off = 0;
  char *base = nullptr;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  }
  template <class T>
  T *at(size_t o) const {
    return reinterpret_cast<T *>(base + o);
  }
};

}  // namespace

namespace skw {
skw_status upload_irradiance(skw_irradiance_map *m) {
  SKW_HIP(upload(&m->d_e_ext, m->e_ext), "upload irradiance");
  SKW_HIP(upload(&m->d_e_int, m->e_int), "upload irradiance");
  return SKW_OK;
}

}  // namespace skw

namespace {

// The irradiance photons' tree, with the normal boxes its nearest lookups prune by.
skw_status build_irr_tree(skw_irradiance_map *m) {
  const uint64_t n = m->normal.size();
  std::vector<uint16_t> dir(n, 0);
  skw_status st = skw_photon_map_create(n, m->pos.data(), m->e_ext.data(), dir.data(), m->normal.data(), 0, &m->map);
  if (st != SKW_OK) return st;
  return photon_map_nbox(m->map, nullptr);
}

// The lookup distance: 0.1 % percentile (index max(size / 1000, 1) of the descending order) of the
// squared distances from every 11th indirect photon to its nearest irradiance photon; infinite when
// that distance is 0 (PhotonLTEIntegrator.cpp:226-244).
skw_status lookup_distance(skw_irradiance_map *m, const KdLayout &ind, hipStream_t st) {
  std::vector<double> qp, qn;
  strided_queries(ind, 11, qp, qn);
  const size_t nq = qp.size() / 3;
  if (nq == 0) return SKW_OK;
  Block b;
  const size_t o_p = b.take(qp.size() * sizeof(double)), o_n = b.take(qn.size() * sizeof(double));
  const size_t o_id = b.take(nq * sizeof(uint32_t));
  SKW_HIP(pool_alloc((void **)&b.base, b.off, st), "lookup distance buffers");
  std::vector<uint32_t> id(nq);
  hipError_t e = hipMemcpyAsync(b.at<double>(o_p), qp.data(), qp.size() * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(b.at<double>(o_n), qn.data(), qn.size() * sizeof(double), hipMemcpyHostToDevice, st);
  skw_status s = SKW_OK;
  if (e == hipSuccess)
    s = skw_photon_nearest(m->map, b.at<double>(o_p), b.at<double>(o_n), nq, inf(), 1, b.at<uint32_t>(o_id), st);
  if (e == hipSuccess && s == SKW_OK)
    e = hipMemcpyAsync(id.data(), b.at<uint32_t>(o_id), nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && s == SKW_OK) e = hipStreamSynchronize(st);
  pool_free(b.base, st);
  if (s != SKW_OK) return s;
  if (e != hipSuccess) return hip_fail(e, "irradiance lookup distance");
  std::vector<double> sq(nq, 0.0);
  for (size_t i = 0; i < nq; ++i) {
    if (id[i] == kMiss) continue;
    const double dx = qp[3 * i] - (double)m->pos[3 * (size_t)id[i]];
    const double dy = qp[3 * i + 1] - (double)m->pos[3 * (size_t)id[i] + 1];
    const double dz = qp[3 * i + 2] - (double)m->pos[3 * (size_t)id[i] + 2];
    sq[i] = dx * dx + dy * dy + dz * dz;
  }
  const size_t index = nq <= 1 ? 0 : std::max(nq / 1000, (size_t)1);
  std::nth_element(sq.begin(), sq.begin() + index, sq.end(), std::greater<double>());
  m->max_dist = sq[index] == 0.0 ? inf() : std::sqrt(sq[index]);
  return SKW_OK;
}

bool device_map(const skw_photon_map *m) { return m && m->on_device; }

}  // namespace

extern "C" {

skw_status skw_irradiance_map_create(const skw_irradiance_map_desc *desc, skw_irradiance_map **out, void *stream) {
  if (!desc || !out) return fail(SKW_INVALID_ARG, "desc and out must not be NULL");
  *out = nullptr;
  if (!desc->direct || !desc->indirect) return fail(SKW_INVALID_ARG, "direct and indirect maps are required");
  if (!device_map(desc->direct) || !device_map(desc->indirect) || (desc->caustic && !device_map(desc->caustic)))
    return fail(SKW_UNSUPPORTED, "photon maps must be device maps");
  if (!(desc->scene_area >= 0.0) || !std::isfinite(desc->scene_area))
    return fail(SKW_INVALID_ARG, "scene_area must be finite and >= 0");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const KdLayout &ind = photon_map_host(desc->indirect);
  if (ind.nodes.size() != desc->indirect->n) return fail(SKW_HIP_ERROR, "photon tree download failed");
  std::unique_ptr<skw_irradiance_map> m(new skw_irradiance_map());
  std::vector<double> pts, nrm;
  strided_queries(ind, 10, pts, nrm);
  const uint64_t n = pts.size() / 3;
  m->pos.resize(3 * n);
  m->normal.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) m->pos[3 * i + c] = ind.nodes[10 * i].p[c];
    m->normal[i] = (uint16_t)(ind.photons[10 * i].codes >> 16);
  }
  m->e_ext.assign(3 * n, 0.0f);
  m->e_int.assign(3 * n, 0.0f);
  if (n == 0) {  // the reference keeps no irradiance map: every lookup finds nothing
    *out = m.release();
    return SKW_OK;
  }
  // Radii holding 100 photons of each map on average over the scene area (:187-196).
  const size_t nd = desc->direct->n, ni = ind.nodes.size();
  const double r_direct = std::sqrt((nd == 0 ? 0.0 : desc->scene_area / (double)nd) * 100.0 * kInvPi);
  const double r_indirect = std::sqrt((ni == 0 ? 0.0 : desc->scene_area / (double)ni) * 100.0 * kInvPi);
  const double r_caustic = std::max(r_direct, r_indirect);
  Block b;
  const size_t o_p = b.take(3 * n * sizeof(double)), o_n = b.take(3 * n * sizeof(double));
  size_t o_e[6];
  for (int j = 0; j < 6; ++j) o_e[j] = b.take(3 * n * sizeof(float));
  const size_t o_sum_e = b.take(3 * n * sizeof(float)), o_sum_i = b.take(3 * n * sizeof(float));
  SKW_HIP(pool_alloc((void **)&b.base, b.off, st), "irradiance map buffers");
  float *e[6];
  for (int j = 0; j < 6; ++j) e[j] = b.at<float>(o_e[j]);
  const double *dp = b.at<double>(o_p), *dn = b.at<double>(o_n);
  hipError_t he = hipMemcpyAsync(b.at<double>(o_p), pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, st);
  if (he == hipSuccess)
    he = hipMemcpyAsync(b.at<double>(o_n), nrm.data(), nrm.size() * sizeof(double), hipMemcpyHostToDevice, st);
  skw_status s = SKW_OK;
  if (he == hipSuccess && desc->caustic)
    s = skw_photon_irradiance(desc->caustic, dp, dn, n, 100, r_caustic, desc->caustic_paths, 0.0, e[0], e[1], nullptr,
                              nullptr, nullptr, nullptr, st);
  if (he == hipSuccess && s == SKW_OK)
    s = skw_photon_irradiance(desc->direct, dp, dn, n, 100, r_direct, desc->direct_paths, 0.0, e[2], e[3], nullptr,
                              nullptr, nullptr, nullptr, st);
  if (he == hipSuccess && s == SKW_OK)
    s = skw_photon_irradiance(desc->indirect, dp, dn, n, 100, r_indirect, desc->indirect_paths, 0.0, e[4], e[5],
                              nullptr, nullptr, nullptr, nullptr, st);
  // E = (caustic + direct) + indirect in float, external and internal sides.
  if (he == hipSuccess && s == SKW_OK)
    he = launch_irr_sum(desc->caustic ? e[0] : nullptr, e[2], e[4], b.at<float>(o_sum_e), 3 * n, st);
  if (he == hipSuccess && s == SKW_OK)
    he = launch_irr_sum(desc->caustic ? e[1] : nullptr, e[3], e[5], b.at<float>(o_sum_i), 3 * n, st);
  if (he == hipSuccess && s == SKW_OK)
    he = hipMemcpyAsync(m->e_ext.data(), b.at<float>(o_sum_e), 3 * n * sizeof(float), hipMemcpyDeviceToHost, st);
  if (he == hipSuccess && s == SKW_OK)
    he = hipMemcpyAsync(m->e_int.data(), b.at<float>(o_sum_i), 3 * n * sizeof(float), hipMemcpyDeviceToHost, st);
  if (he == hipSuccess && s == SKW_OK) he = hipStreamSynchronize(st);
  pool_free(b.base, st);
  if (s != SKW_OK) return s;
  if (he != hipSuccess) return hip_fail(he, "irradiance estimates");
  if ((s = build_irr_tree(m.get())) != SKW_OK) return s;
  if ((s = upload_irradiance(m.get())) != SKW_OK) return s;
  if ((s = lookup_distance(m.get(), ind, st)) != SKW_OK) return s;
  *out = m.release();
  return SKW_OK;
}

skw_status skw_irradiance_map_create_from(uint64_t n, const float *pos, const uint16_t *normal, const float *e_ext,
                                          const float *e_int, double max_lookup_dist, skw_irradiance_map **out) {
  if (!out) return fail(SKW_INVALID_ARG, "out must not be NULL");
  *out = nullptr;
  if (n > 0 && (!pos || !normal || !e_ext || !e_int)) return fail(SKW_INVALID_ARG, "arrays must not be NULL");
  if (!(max_lookup_dist >= 0.0)) return fail(SKW_INVALID_ARG, "max_lookup_dist must be >= 0");
  if (device_count() == 0) return fail(SKW_NO_DEVICE, "no HIP device is visible");
  std::unique_ptr<skw_irradiance_map> m(new skw_irradiance_map());
  m->pos.assign(pos, pos + 3 * n);
  m->normal.assign(normal, normal + n);
  m->e_ext.assign(e_ext, e_ext + 3 * n);
  m->e_int.assign(e_int, e_int + 3 * n);
  m->max_dist = max_lookup_dist;
  if (n > 0) {
    skw_status s = build_irr_tree(m.get());
    if (s == SKW_OK) s = upload_irradiance(m.get());
    if (s != SKW_OK) return s;
  }
  *out = m.release();
  return SKW_OK;
}

skw_status skw_irradiance_map_destroy(skw_irradiance_map *map) {
  delete map;
  return SKW_OK;
}

skw_status skw_irradiance_map_info(const skw_irradiance_map *map, uint64_t *n, double *max_lookup_dist) {
  if (!map) return fail(SKW_INVALID_ARG, "map must not be NULL");
  if (n) *n = map->normal.size();
  if (max_lookup_dist) *max_lookup_dist = map->max_dist;
  return SKW_OK;
}

skw_status skw_irradiance_map_export(const skw_irradiance_map *map, float *pos, uint16_t *normal, float *e_ext,
                                     float *e_int) {
  if (!map) return fail(SKW_INVALID_ARG, "map must not be NULL");
  if (pos) std::copy(map->pos.begin(), map->pos.end(), pos);
  if (normal) std::copy(map->normal.begin(), map->normal.end(), normal);
  if (e_ext) std::copy(map->e_ext.begin(), map->e_ext.end(), e_ext);
  if (e_int) std::copy(map->e_int.begin(), map->e_int.end(), e_int);
  return SKW_OK;
}

skw_status skw_dl_final_gather(const skw_scene *scene, const skw_dl_scene *dl, const skw_irradiance_map *irr,
                               const skw_photon_map *indirect, double scene_area, const skw_camera *cam,
                               const skw_fg_params *prm, float *d_rgb, double *d_radiance, uint64_t *rays_out,
                               void *stream) {
  if (!scene || !dl || !irr || !cam || !prm) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (indirect && !indirect->on_device) return fail(SKW_UNSUPPORTED, "photon map was created host-only");
  if (prm->width == 0 || prm->height == 0 || prm->spp == 0) return fail(SKW_INVALID_ARG, "empty image");
  if (prm->gather_samples == 0 || prm->gather_samples > kMaxGather)
    return fail(SKW_INVALID_ARG, "gather_samples must be in [1, 676]");
  if (!(scene_area >= 0.0) || !std::isfinite(scene_area))
    return fail(SKW_INVALID_ARG, "scene_area must be finite and >= 0");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const uint64_t npix = (uint64_t)prm->width * prm->height;
  const uint32_t gq = (uint32_t)std::ceil(std::sqrt((double)prm->gather_samples));
  const uint32_t gs = gq * gq;
  const uint64_t n_ind = indirect ? indirect->n : 0;
  // PhotonLTEIntegrator.cpp:460-462: the radius holding 32 indirect photons on average.
  const double lookup_r = n_ind > 0 ? std::sqrt((scene_area / (double)n_ind) * kFgK * kInvPi) : 0.0;
  // Launches: a pixel batch (every slot holds up to 2 gs gather rays) and, when one batch holds every
  // pixel, a power-of-two group of sample passes (at most 16, dividing spp): slot i is pixel
  // i >> ppl_log2 of pass pass + (i & (passes - 1)), as in the direct-lighting frame. Neither
  // changes any value (the sample stream depends on the pixel and the sample only, and the film
  // adds a pixel's samples in pass order); larger launches mean fewer launch tails.
  // Bytes held per sample slot and per gather ray (the buffers laid out below).
  const uint64_t per_slot = sizeof(skw_ray) + sizeof(skw_hit) + (3 + 3 + 24 + 3 + 3 + 4) * sizeof(double) +
                            kFgK * sizeof(uint32_t) + sizeof(uint32_t);
  const uint64_t per_ray = sizeof(skw_ray) + sizeof(skw_hit) + (3 + 3 + 3 + 3) * sizeof(double) + 1 +
                           6 * sizeof(uint32_t) + 16;  // + radix-sort workspace, about 8 B per key
  const uint64_t slot_bytes = per_slot + 2ull * gs * per_ray;
  const int rays_log2 = std::min(28, std::max(10, env_int("SKW_FG_GATHER_RAYS_LOG2", kMaxGatherRaysLog2)));
  const uint64_t rays_cap = uint64_t(1) << rays_log2;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  uint64_t batch = std::max<uint64_t>(64, rays_cap / (2ull * gs));
  batch = std::min(batch, npix);
  while (batch > 64 && batch * slot_bytes > free_b / 4) batch = (batch + 1) / 2;
  uint32_t ppl_log2 = 0;
  if (batch == npix)
    while (ppl_log2 < kMaxPassesPerLaunchLog2 && prm->spp % (2u << ppl_log2) == 0 &&
           (npix << (ppl_log2 + 1)) * 2ull * gs <= rays_cap && (npix << (ppl_log2 + 1)) * slot_bytes <= free_b / 4)
      ++ppl_log2;
  const uint64_t slots = batch << ppl_log2;
  const uint64_t max_rays = (slots + kFgBlock - 1) / kFgBlock * kFgBlock * 2ull * gs;

  const DevMat *mats = nullptr;
  const double *rad = nullptr;
  dl_tables(dl, &mats, &rad);
  FgArgs f;
  std::memset(&f, 0, sizeof(f));
  DlArgs &a = f.a;
  a.geom = scene->geom_args();
  a.mats = mats;
  a.mats_e = mats;
  a.rad_e = rad;
  a.rad_base = rad;
  a.n_lights = 0;  // no light CDFs: the final gather samples no lights
  for (int i = 0; i < 3; ++i) a.cam_o[i] = cam->origin[i];
  for (int i = 0; i < 9; ++i) a.cam_r[i] = cam->rot[i];
  a.x_tan = cam->x_tan;
  a.y_tan = cam->y_tan;
  a.width = prm->width;
  a.height = prm->height;
  a.spp = prm->spp;
  a.seed = prm->seed;
  f.gs = gs;
  f.gs_sqrt = gq;
  f.max_rays = max_rays;
  f.cdir = n_ind > 0 ? indirect->d_cdir : nullptr;
  f.e_ext = irr->d_e_ext;
  f.e_int = irr->d_e_int;
  f.radiance = d_radiance;

  Block b;
  const size_t o_prays = b.take(slots * sizeof(skw_ray)), o_phits = b.take(slots * sizeof(skw_hit));
  const size_t o_le = b.take(slots * 3 * sizeof(double)), o_ns = b.take(slots * 3 * sizeof(double));
  const size_t o_cdf = b.take(sizeof(double)), o_point = b.take(slots * 24 * sizeof(double));
  const size_t o_spts = b.take(slots * 3 * sizeof(double)), o_snrm = b.take(slots * 3 * sizeof(double));
  const size_t o_nids = b.take(slots * kFgK * sizeof(uint32_t)), o_ncnt = b.take(slots * sizeof(uint32_t));
  const size_t o_slot_l = ppl_log2 > 0 ? b.take(slots * 4 * sizeof(double)) : 0;
  const size_t o_rays = b.take(max_rays * sizeof(skw_ray)), o_w = b.take(max_rays * 3 * sizeof(double));
  const size_t o_hits = b.take(max_rays * sizeof(skw_hit));
  const size_t o_qp = b.take(max_rays * 3 * sizeof(double)), o_qn = b.take(max_rays * 3 * sizeof(double));
  const size_t o_ts = b.take(max_rays * 3 * sizeof(double)), o_side = b.take(max_rays);
  const size_t o_irr = b.take(max_rays * sizeof(uint32_t));
  const size_t o_plastic = b.take(max_rays * sizeof(uint32_t)), o_pcount = b.take(sizeof(uint64_t));
  // irradiance lookups in spatial order: cell keys of the lookup points, radix-sorted with the
  // gather-ray indices
  const size_t o_keys = b.take(max_rays * sizeof(uint32_t)), o_keys2 = b.take(max_rays * sizeof(uint32_t));
  const size_t o_ord = b.take(max_rays * sizeof(uint32_t)), o_ord2 = b.take(max_rays * sizeof(uint32_t));
  size_t sort_bytes = 0;
  const int mbits = kMortonBits;
  const int kKeyLo = 32 - 3 * mbits, kKeyHi = 32;  // the key bits of the Morton cells
  SKW_HIP(sort_pairs_u32(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, max_rays, kKeyLo, kKeyHi, st),
          "sort workspace size");
  const size_t o_sort = b.take(sort_bytes);
  const uint32_t n_mats = dl_material_count(dl);
  const size_t o_quad = b.take((10 + 10 * (size_t)n_mats) * sizeof(double));
  const size_t o_counts = b.take(2 * sizeof(uint64_t)), o_totals = b.take(9 * sizeof(uint64_t));
  const size_t o_film = b.take(npix * 3 * sizeof(double));
  SKW_HIP(pool_alloc((void **)&b.base, b.off, st), "final gather buffers");
  skw_status rs0 = SKW_OK;
  a.prim_rays = b.at<skw_ray>(o_prays);
  a.prim_hits = b.at<skw_hit>(o_phits);
  a.le = b.at<double>(o_le);
  a.ns = b.at<double>(o_ns);
  a.cdf = b.at<double>(o_cdf);
  a.point = b.at<double>(o_point);
  a.counts = b.at<uint64_t>(o_counts);
  a.totals = b.at<uint64_t>(o_totals);
  a.film = b.at<double>(o_film);
  a.ppl_log2 = ppl_log2;
  a.slot_L = ppl_log2 > 0 ? b.at<double>(o_slot_l) : nullptr;
  f.quad = b.at<double>(o_quad);
  f.morton_bits = (uint32_t)mbits;
  f.plastic = b.at<uint32_t>(o_plastic);
  f.plastic_count = b.at<uint64_t>(o_pcount);
  f.n_mats = n_mats;
  f.s_pts = b.at<double>(o_spts);
  f.s_nrm = b.at<double>(o_snrm);
  f.near_ids = b.at<uint32_t>(o_nids);
  f.near_count = n_ind > 0 ? b.at<uint32_t>(o_ncnt) : nullptr;
  f.rays = b.at<skw_ray>(o_rays);
  f.w = b.at<double>(o_w);
  f.count = a.counts;
  f.hits = b.at<skw_hit>(o_hits);
  f.q_pts = b.at<double>(o_qp);
  f.q_nrm = b.at<double>(o_qn);
  f.ts = b.at<double>(o_ts);
  f.side = b.at<uint8_t>(o_side);
  f.irr_id = b.at<uint32_t>(o_irr);
  uint32_t *irr_id = b.at<uint32_t>(o_irr);
  f.keys = b.at<uint32_t>(o_keys);
  f.order = b.at<uint32_t>(o_ord);
  uint32_t *keys2 = b.at<uint32_t>(o_keys2), *order2 = b.at<uint32_t>(o_ord2);
  double box[6];
  if ((rs0 = skw_scene_bounds(scene, box)) != SKW_OK) {
    pool_free(b.base, st);
    return rs0;
  }
  bool box_finite = true;
  for (int c = 0; c < 3; ++c) {
    const double ext = box[3 + c] - box[c];
    f.lo[c] = box[c];
    f.cell[c] = ext > 0.0 && std::isfinite(ext) ? (double)(1u << mbits) / ext : 0.0;
    box_finite = box_finite && std::isfinite(box[c]);
  }
  // The ray of a sample that casts none: from below the scene box's lower corner, pointing away
  // (finite inverse direction components, so its wave keeps the NaN-free slab test); with
  // tmax < tmin it can accept no triangle wherever it starts.
  for (int c = 0; c < 3; ++c) {
    f.hole_o[c] = box_finite ? box[c] - 1.0 : 0.0;
    f.hole_d[c] = -0.57735026918962584;
  }
  uint32_t *near_ids = b.at<uint32_t>(o_nids), *near_count = b.at<uint32_t>(o_ncnt);

  hipError_t e = launch_dl_clear(a, st);
  if (e == hipSuccess) e = launch_fg_quad(f, st);
  skw_status rs = SKW_OK;
  for (uint32_t s = 0; s < prm->spp && e == hipSuccess && rs == SKW_OK; s += 1u << ppl_log2) {
    for (uint64_t p0 = 0; p0 < npix && e == hipSuccess && rs == SKW_OK; p0 += batch) {
      a.pass = s;
      a.pix0 = p0;
      a.batch = std::min(batch, npix - p0) << ppl_log2;
      f.slots_padded = (a.batch + kFgBlock - 1) / kFgBlock * kFgBlock;
      e = launch_dl_camera(a, st);  // also zeroes the gather-ray counter
      if (e != hipSuccess) break;
      rs = enqueue_trace(scene, a.prim_rays, a.batch, nullptr, a.prim_hits, nullptr, st, false, nullptr, nullptr,
                         false);
      if (rs != SKW_OK) break;
      e = launch_dl_point(a, st);
      if (e == hipSuccess) e = launch_fg_lookup_in(f, st);
      if (e != hipSuccess) break;
      if (n_ind > 0) {
        rs = photon_knn_codes(indirect, f.s_pts, f.s_nrm, a.batch, kFgK, lookup_r, near_ids, near_count, st);
        if (rs != SKW_OK) break;
      }
      e = launch_fg_samples(f, st);
      if (e != hipSuccess) break;
      rs = enqueue_trace(scene, f.rays, f.slots_padded * 2ull * gs, nullptr, f.hits, nullptr, st, false, nullptr, nullptr,
                         false);
      if (rs != SKW_OK) break;
      e = launch_fg_hit(f, st);
      if (e != hipSuccess) break;
      if (irr->map) {
        const uint64_t nq = f.slots_padded * 2ull * gs;
        size_t bytes = sort_bytes;
        e = sort_pairs_u32(b.base + o_sort, bytes, f.keys, keys2, f.order, order2, nq, kKeyLo, kKeyHi, st);
        if (e != hipSuccess) break;
        rs = photon_nearest_ordered(irr->map, f.q_pts, f.q_nrm, nq, irr->max_dist, order2, irr_id, st);
        if (rs != SKW_OK) break;
      } else {
        e = hipMemsetAsync(irr_id, 0xFF, max_rays * sizeof(uint32_t), st);
        if (e != hipSuccess) break;
      }
      e = launch_fg_resolve(f, st);
      if (e == hipSuccess && a.slot_L) e = launch_dl_accum(a, st);
    }
  }
  if (e == hipSuccess && rs == SKW_OK && d_rgb) {
    a.pix0 = 0;
    a.ppl_log2 = 0;
    a.batch = npix;
    e = launch_dl_finish(a, d_rgb, st);
  }
  uint64_t totals[2] = {0, 0};
  if (e == hipSuccess && rs == SKW_OK && rays_out) {
    e = hipMemcpyAsync(totals, a.totals, sizeof(totals), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  pool_free(b.base, st);
  if (rs != SKW_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "final gather frame");
  if (rays_out) {
    rays_out[0] = totals[0];
    rays_out[1] = totals[1];
  }
  return SKW_OK;
}

}  // extern "C"
