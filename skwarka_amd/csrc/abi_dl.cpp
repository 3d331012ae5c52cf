// C ABI of the direct-lighting render (skw_dl_*): host tables for lights and materials and the
// per-frame wavefront (camera -> closest hit -> shade -> occlusion + BSDF rays -> resolve).
//
// Light tables follow DiffuseAreaLightSource's constructor (Raytracer/LightSources/
// DiffuseAreaLightSource.cpp:20-55): per-triangle float areas accumulated into a double CDF and
// normalised, float triangle normals (TriangleMesh::GetTriangleNormal, TriangleMesh.h:219-232);
// intensities follow IrradianceLightsSamplingStrategy's constructor
// (IrradianceLightsSamplingStrategy.cpp:19-31): luminance of Power() = (pi * area) * radiance over
// pi times the area of the light mesh's bounding box. Material parameters are clamped as
// MatteMaterial / PlasticMaterial do (MatteMaterial.cpp:33-34, PlasticMaterial.cpp:37-45).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace {

constexpr double kPi = 3.14159265358979323846;
// Sample passes per launch of a direct-lighting frame (SKW_DL_PASSES_PER_LAUNCH overrides).
constexpr int kDlPassesPerLaunch = 16;
constexpr int kDlMaxPassesPerLaunch = 16;
constexpr uint32_t kMaxLights = 16;
const double kLumW[3] = {0x1.b38293da589a6p-5, 0x1.2b993cc221566p-1, 0x1.725d3400723f6p-2};

double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

}  // namespace

struct skw_dl_scene {
  std::vector<DevMat> mats;
  std::vector<DevLight> lights;
  std::vector<DevLightTri> ltris;
  std::vector<double> rad;  // 3 per light
  uint32_t n_groups = 0;
  DevMat *d_mats = nullptr;
  DevLight *d_lights = nullptr;
  DevLightTri *d_ltris = nullptr;
  double *d_rad = nullptr;
  ~skw_dl_scene() {
    void *ptrs[] = {d_mats, d_lights, d_ltris, d_rad};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
};

// Per-sample derivatives of one replayed frame (see skw_dl_render_tape).
struct skw_dl_tape {
  double *d_j = nullptr;
  uint32_t *d_group = nullptr;
  uint32_t stride = 0, n_groups = 0, n_lights = 0, spp = 0, width = 0, height = 0;
  uint64_t batch = 0, n_samples = 0;
  ~skw_dl_tape() {
    if (d_j) (void)hipFree(d_j);
    if (d_group) (void)hipFree(d_group);
  }
};

namespace {

// What a frame produces: a float image, a double image, (adj set) a gradient, or (tape set) a
// double image and the tape.
struct FrameOut {
  double *sample_L = nullptr;  // per-sample Le + direct lighting (no film, no image)
  float *rgb = nullptr;
  double *img = nullptr;
  const double *adj = nullptr;
  double *grad = nullptr;
  skw_dl_tape *tape = nullptr;
};

// Parameter tables of one call, placed at the front of the frame's device block.
struct ParamPlan {
  bool active = false;
  bool alias = false;  // theta_eval equals theta_sample: one material table
  const double *theta_s = nullptr, *theta_e = nullptr;
};

skw_status check_frame_args(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                            const skw_dl_params *prm) {
  if (!scene || !dl || !cam || !prm) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (prm->width == 0 || prm->height == 0 || prm->spp == 0) return fail(SKW_INVALID_ARG, "empty image");
  if (prm->light_samples == 0 || prm->light_samples > 8 || prm->bsdf_samples == 0 || prm->bsdf_samples > 8)
    return fail(SKW_INVALID_ARG, "light_samples and bsdf_samples must be in [1, 8]");
  // shadow / BSDF ray slots of a launch are 32-bit indices
  if ((uint64_t)prm->width * prm->height * std::max(prm->light_samples, prm->bsdf_samples) > (uint64_t)INT32_MAX)
    return fail(SKW_INVALID_ARG, "image too large: width x height x samples per hit must stay below 2^31");
  return SKW_OK;
}

// The wavefront of one frame: per launch of P sample passes, camera -> closest hit -> shade ->
// occlusion and BSDF rays (ray counts read on the device) -> resolve; then the image (or nothing,
// for the adjoint, whose resolve accumulates the gradient). A launch covers P passes of every
// pixel (P a power of two dividing spp, at most kDlMaxPassesPerLaunch): fewer, larger trace
// launches with shorter drains, and the P samples of a pixel in adjacent slots (coherent waves).
// Every sample's values are those of its own one-pass launch, and the film adds a pixel's samples
// in pass order, so the image does not depend on P.
// prof (optional): HIP-event time per stage summed over the passes; count: also count node and
// triangle tests per ray class (the walk's counting variant).
skw_status run_frame(const skw_scene *scene, const skw_dl_scene *dl, const ParamPlan &pp, const skw_camera *cam,
                     const skw_dl_params *prm, const FrameOut &out, uint64_t *rays_out, hipStream_t st,
                     skw_dl_profile *prof = nullptr, bool count = false) {
  const uint64_t npix = (uint64_t)prm->width * prm->height;
  const uint32_t nls = prm->light_samples, nbs = prm->bsdf_samples, nl = (uint32_t)dl->lights.size();
  const uint32_t np = (uint32_t)dl->mats.size();
  const bool adj = out.adj != nullptr;
  // Passes per launch: the largest power of two up to the setting that divides spp and keeps the
  // frame buffer (per_slot bytes per sample slot, the per-slot arrays laid out below) within a
  // quarter of the device's free memory.
  uint32_t ppl = (uint32_t)std::min(kDlMaxPassesPerLaunch, std::max(1, env_int("SKW_DL_PASSES_PER_LAUNCH",
                                                                                  kDlPassesPerLaunch)));
  uint32_t ppl_log2 = 0;
  while ((2u << ppl_log2) <= ppl) ++ppl_log2;
  while (ppl_log2 > 0 && prm->spp % (1u << ppl_log2) != 0) --ppl_log2;
  const uint64_t per_slot = sizeof(skw_ray) + sizeof(skw_hit) + (3 + (nl ? nl : 1) + 3 + 24 + 4) * sizeof(double) +
                            nls * (sizeof(skw_ray) + sizeof(int32_t) + 3 * sizeof(double) + 1) +
                            nbs * (sizeof(skw_ray) + sizeof(int32_t) + 8 * sizeof(double) + sizeof(skw_hit)) +
                            ((adj || out.tape) ? nls * 6 * sizeof(double) + sizeof(int32_t) : 0);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  while (ppl_log2 > 0 && (npix * (1ull << ppl_log2) * per_slot > free_b / 4 ||
                          npix * (1ull << ppl_log2) * std::max(nls, nbs) > (uint64_t)INT32_MAX))
    --ppl_log2;
  const uint32_t ppl_eff = 1u << ppl_log2;
  const uint64_t batch = npix * ppl_eff;
  DlArgs a;
  std::memset(&a, 0, sizeof(a));
  a.geom = scene->geom_args();
  a.ltris = dl->d_ltris;
  a.n_lights = nl;
  a.n_groups = dl->n_groups;
  for (int i = 0; i < 3; ++i) a.cam_o[i] = cam->origin[i];
  for (int i = 0; i < 9; ++i) a.cam_r[i] = cam->rot[i];
  a.x_tan = cam->x_tan;
  a.y_tan = cam->y_tan;
  a.width = prm->width;
  a.height = prm->height;
  a.spp = prm->spp;
  a.nls = nls;
  a.nbs = nbs;
  a.seed = prm->seed;
  a.pix0 = 0;
  a.batch = batch;
  a.ppl_log2 = ppl_log2;
  // One device block holds every per-frame buffer.
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  size_t o_mats_s = 0, o_mats_e = 0, o_lights = 0, o_rad_e = 0, o_rad_b = 0;
  if (pp.active) {
    o_mats_s = take(np * sizeof(DevMat));
    if (!pp.alias) o_mats_e = take(np * sizeof(DevMat));
    o_lights = take(nl * sizeof(DevLight));
    o_rad_e = take(nl * 3 * sizeof(double));
    o_rad_b = take(nl * 3 * sizeof(double));
  }
  size_t o_prim_rays = take(batch * sizeof(skw_ray)), o_prim_hits = take(batch * sizeof(skw_hit));
  size_t o_le = take(batch * 3 * sizeof(double)), o_cdf = take(batch * (nl ? nl : 1) * sizeof(double));
  size_t o_ns = take(batch * 3 * sizeof(double)), o_point = take(batch * 24 * sizeof(double));
  size_t o_srays = take(batch * nls * sizeof(skw_ray)), o_sidx = take(batch * nls * sizeof(int32_t));
  size_t o_sc = take(batch * nls * 3 * sizeof(double)), o_occ = take(batch * nls);
  size_t o_brays = take(batch * nbs * sizeof(skw_ray)), o_bidx = take(batch * nbs * sizeof(int32_t));
  size_t o_brec = take(batch * nbs * 8 * sizeof(double)), o_bhits = take(batch * nbs * sizeof(skw_hit));
  size_t o_counts = take(2 * sizeof(uint64_t)), o_totals = take(9 * sizeof(uint64_t));
  size_t o_film = take(npix * 3 * sizeof(double));
  size_t o_saux = 0, o_caml = 0, o_slot_l = 0;
  const bool slot_film = ppl_eff > 1 && !adj && !out.sample_L;
  if (slot_film) o_slot_l = take(batch * 4 * sizeof(double));
  if (adj || out.tape) {
    o_saux = take(batch * nls * 6 * sizeof(double));
    o_caml = take(batch * sizeof(int32_t));
  }
  char *base = nullptr;
  SKW_HIP(pool_alloc((void **)&base, off, st), "frame buffer allocation");
  hipError_t e = hipSuccess;
  if (pp.active) {
    DlParamArgs pa;
    std::memset(&pa, 0, sizeof(pa));
    pa.base_mats = dl->d_mats;
    pa.base_lights = dl->d_lights;
    pa.n_prims = np;
    pa.n_lights = nl;
    pa.n_groups = dl->n_groups;
    pa.mats_s = (DevMat *)(base + o_mats_s);
    pa.mats_e = pp.alias ? pa.mats_s : (DevMat *)(base + o_mats_e);
    pa.lights_s = (DevLight *)(base + o_lights);
    pa.rad_e = (double *)(base + o_rad_e);
    pa.rad_base = (double *)(base + o_rad_b);
    const uint32_t n_params = 2 * dl->n_groups + nl;
    for (uint32_t j = 0; j < n_params; ++j) {
      pa.theta_s[j] = pp.theta_s[j];
      pa.theta_e[j] = pp.theta_e[j];
    }
    e = launch_dl_params(pa, st);
    a.mats = pa.mats_s;
    a.mats_e = pa.mats_e;
    a.lights = pa.lights_s;
    a.rad_e = pa.rad_e;
    a.rad_base = pa.rad_base;
  } else {
    a.mats = dl->d_mats;
    a.mats_e = dl->d_mats;
    a.lights = dl->d_lights;
    a.rad_e = dl->d_rad;
    a.rad_base = dl->d_rad;
  }
  a.prim_rays = (skw_ray *)(base + o_prim_rays);
  a.prim_hits = (skw_hit *)(base + o_prim_hits);
  a.le = (double *)(base + o_le);
  a.cdf = (double *)(base + o_cdf);
  a.ns = (double *)(base + o_ns);
  a.point = (double *)(base + o_point);
  a.shadow_rays = (skw_ray *)(base + o_srays);
  a.shadow_idx = (int32_t *)(base + o_sidx);
  a.shadow_c = (double *)(base + o_sc);
  a.occluded = (uint8_t *)(base + o_occ);
  a.bsdf_rays = (skw_ray *)(base + o_brays);
  a.bsdf_idx = (int32_t *)(base + o_bidx);
  a.bsdf_rec = (double *)(base + o_brec);
  a.bsdf_hits = (skw_hit *)(base + o_bhits);
  a.counts = (uint64_t *)(base + o_counts);
  a.totals = (uint64_t *)(base + o_totals);
  a.film = (double *)(base + o_film);
  if (slot_film) a.slot_L = (double *)(base + o_slot_l);
  a.sample_L = out.sample_L;
  if (adj || out.tape) {
    a.adj = out.adj;
    a.grad = out.grad;
    a.shadow_aux = (double *)(base + o_saux);
    a.cam_light = (int32_t *)(base + o_caml);
  }
  if (out.tape) {
    a.tape_j = out.tape->d_j;
    a.tape_group = out.tape->d_group;
    a.tape_stride = out.tape->stride;
  }
  skw_status rs = SKW_OK;
  // stage boundaries of every pass: before camera, after camera / primary trace / shade /
  // occlusion trace / BSDF trace / resolve
  std::vector<hipEvent_t> ev;
  if (prof) {
    ev.assign(6 * (size_t)(prm->spp / ppl_eff) + 1, nullptr);
    for (auto &x : ev)
      if (e == hipSuccess) e = hipEventCreate(&x);
  }
  size_t ev_next = 0;
  auto mark = [&]() {
    if (prof && e == hipSuccess) e = hipEventRecord(ev[ev_next++], st);
  };
  uint64_t *sums = count ? a.totals + 3 : nullptr;
  // The occlusion and BSDF traces of a launch are independent. When a launch covers one pass, the
  // BSDF trace runs on a second stream beside the occlusion trace, so one trace's tail (its last
  // long walks) overlaps the other. Launches of several passes have tails short against their
  // length, and two persistent grids sharing the chip cost more than the overlap saves (c2 frame
  // 64.2 ms with the second stream, 63.0 ms without). A profiled frame keeps one stream, so that
  // its events time each stage alone. SKW_DL_SIDE_STREAM=0 / 1 forces the choice.
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  if (!prof && env_int("SKW_DL_SIDE_STREAM", ppl_eff == 1 ? 1 : 0) != 0) {
    if (e == hipSuccess) e = side_stream(&st2);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_join, hipEventDisableTiming);
  }
  if (e == hipSuccess) e = launch_dl_clear(a, st);
  if (e == hipSuccess && adj) e = launch_dl_clear_grad(out.grad, 2 * dl->n_groups + nl, st);
  mark();
  for (uint32_t s = 0; s < prm->spp && e == hipSuccess && rs == SKW_OK; s += ppl_eff) {
    a.pass = s;
    e = launch_dl_camera(a, st);
    if (e != hipSuccess) break;
    mark();
    rs = enqueue_trace(scene, a.prim_rays, batch, nullptr, a.prim_hits, nullptr, st, false, nullptr, nullptr, false,
                       sums);
    if (rs != SKW_OK) break;
    mark();
    e = launch_dl_shade(a, st);
    if (e != hipSuccess) break;
    mark();
    if (st2) {
      e = hipEventRecord(ev_fork, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(st2, ev_fork, 0);
      if (e != hipSuccess) break;
      rs = enqueue_trace(scene, a.bsdf_rays, batch * nbs, a.counts + 1, a.bsdf_hits, nullptr, st2, false, nullptr,
                         nullptr, false, sums ? sums + 4 : nullptr);
      if (rs != SKW_OK) break;
      e = hipEventRecord(ev_join, st2);
      if (e != hipSuccess) break;
    }
    rs = enqueue_trace(scene, a.shadow_rays, batch * nls, a.counts, nullptr, a.occluded, st, true, nullptr, nullptr,
                       false, sums ? sums + 2 : nullptr);
    if (rs != SKW_OK) break;
    mark();
    if (st2) {
      e = hipStreamWaitEvent(st, ev_join, 0);
      if (e != hipSuccess) break;
    } else {
      rs = enqueue_trace(scene, a.bsdf_rays, batch * nbs, a.counts + 1, a.bsdf_hits, nullptr, st, false, nullptr,
                         nullptr, false, sums ? sums + 4 : nullptr);
      if (rs != SKW_OK) break;
    }
    mark();
    e = launch_dl_resolve(a, st);
    if (e == hipSuccess && a.slot_L) e = launch_dl_accum(a, st);
    mark();
  }
  if (e == hipSuccess && rs == SKW_OK) {
    if (out.rgb) e = launch_dl_finish(a, out.rgb, st);
    else if (out.img) e = launch_dl_finish_f64(a, out.img, st);
  }
  uint64_t totals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (e == hipSuccess && rs == SKW_OK && (rays_out || prof)) {
    e = hipMemcpyAsync(totals, a.totals, sizeof(totals), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  if (prof && e == hipSuccess && rs == SKW_OK) {
    std::memset(prof, 0, sizeof(*prof));
    for (size_t i = 0; i + 1 < ev.size() && e == hipSuccess; ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      prof->stage_ms[i % 6] += ms;
    }
    for (int c = 0; c < 3; ++c) {
      prof->rays[c] = totals[c];
      prof->node_tests[c] = totals[3 + 2 * c];
      prof->tri_tests[c] = totals[4 + 2 * c];
    }
    prof->launches = prm->spp / ppl_eff;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (st2 && (rs != SKW_OK || e != hipSuccess)) (void)hipStreamSynchronize(st2);  // nothing left reading `base`
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  pool_free(base, st);
  if (rs != SKW_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "direct lighting frame");
  if (rays_out)
    for (int i = 0; i < 3; ++i) rays_out[i] = totals[i];
  return SKW_OK;
}

skw_status check_theta(const skw_dl_scene *dl, const double *theta) {
  if (!theta) return fail(SKW_INVALID_ARG, "theta must not be NULL");
  const uint32_t n = 2 * dl->n_groups + (uint32_t)dl->lights.size();
  for (uint32_t j = 0; j < n; ++j)
    if (!std::isfinite(theta[j])) return fail(SKW_INVALID_ARG, "theta holds a non-finite value");
  return SKW_OK;
}

}  // namespace

namespace skw {
void dl_tables(const skw_dl_scene *dl, const DevMat **mats, const double **rad) {
  *mats = dl->d_mats;
  *rad = dl->d_rad;
}
uint32_t dl_material_count(const skw_dl_scene *dl) { return (uint32_t)dl->mats.size(); }
void dl_lights(const skw_dl_scene *dl, const std::vector<DevLight> **host, const DevLight **d_lights,
               const DevLightTri **d_ltris) {
  *host = &dl->lights;
  *d_lights = dl->d_lights;
  *d_ltris = dl->d_ltris;
}
}  // namespace skw

extern "C" {

skw_status skw_dl_scene_create(const skw_scene *scene, const skw_dl_desc *desc, skw_dl_scene **out) {
  if (!scene || !desc || !out) return fail(SKW_INVALID_ARG, "scene, desc and out must not be NULL");
  *out = nullptr;
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  const size_t np = scene->host.prims.size();
  if (np > 0 && (!desc->mat_type || !desc->kd || !desc->ks || !desc->roughness))
    return fail(SKW_INVALID_ARG, "material arrays must not be NULL");
  if (desc->n_lights > kMaxLights) return fail(SKW_UNSUPPORTED, "at most 16 area lights");
  if (desc->n_lights > 0 && (!desc->light_prim || !desc->light_radiance))
    return fail(SKW_INVALID_ARG, "light arrays must not be NULL");
  if (desc->n_groups > kDlMaxGroups) return fail(SKW_UNSUPPORTED, "at most 64 parameter groups");
  if (desc->n_groups > 0 && np > 0 && !desc->prim_group)
    return fail(SKW_INVALID_ARG, "prim_group must not be NULL when n_groups > 0");
  std::unique_ptr<skw_dl_scene> d(new skw_dl_scene());
  d->n_groups = desc->n_groups;
  d->mats.resize(np);
  for (size_t p = 0; p < np; ++p) {
    DevMat &m = d->mats[p];
    std::memset(&m, 0, sizeof(m));
    m.group = kMiss;
    if (desc->n_groups > 0) {
      uint32_t g = desc->prim_group[p];
      if (g != SKW_NO_GROUP && g >= desc->n_groups) return fail(SKW_INVALID_ARG, "prim_group out of range");
      m.group = g == SKW_NO_GROUP ? kMiss : g;
    }
    if (desc->mat_type[p] > 1) return fail(SKW_INVALID_ARG, "material type must be 0 (matte) or 1 (plastic)");
    m.type = desc->mat_type[p];
    for (int c = 0; c < 3; ++c) {
      m.kd[c] = clamp01(desc->kd[3 * p + c]);
      m.ks[c] = clamp01(desc->ks[3 * p + c]);
    }
    double r = desc->roughness[p];
    r = r < 0.001 ? 0.001 : (r > 1.0 ? 1.0 : r);
    double e = 1.0 / r;  // BlinnDistribution(1/roughness), exponent clamped to [0, 1000]
    m.exponent = e < 0.0 ? 0.0 : (e > 1000.0 ? 1000.0 : e);
    m.light = kMiss;
  }
  for (uint32_t l = 0; l < desc->n_lights; ++l) {
    uint32_t prim = desc->light_prim[l];
    if (prim >= np) return fail(SKW_INVALID_ARG, "light prim out of range");
    const HostPrim &hp = scene->host.prims[prim];
    if (!hp.identity) return fail(SKW_UNSUPPORTED, "area light primitives must have identity transforms");
    if (d->mats[prim].light != kMiss) return fail(SKW_INVALID_ARG, "a primitive carries two lights");
    d->mats[prim].light = l;
    const HostMesh &mesh = scene->host.meshes[hp.mesh];
    DevLight L;
    for (int c = 0; c < 3; ++c) L.radiance[c] = desc->light_radiance[3 * l + c];
    L.tri_off = (uint32_t)d->ltris.size();
    L.tri_count = (uint32_t)(mesh.tris.size() / 3);
    float fmn[3] = {INFINITY, INFINITY, INFINITY}, fmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    double acc = 0.0;
    size_t first = d->ltris.size();
    for (uint32_t t = 0; t < L.tri_count; ++t) {
      const float *v[3];
      for (int k = 0; k < 3; ++k) v[k] = &mesh.verts[3 * mesh.tris[3 * t + k]];
      // Triangle3D<float>::GetArea: float cross product and length
      float e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
      float e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
      float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
      float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
      acc = (t == 0) ? (double)area : (double)area + acc;
      // GetTriangleNormal: float cross of (v1-v0, v2-v0) or (v2-v0, v1-v0), Normalized() in float
      float nx, ny, nz;
      if (mesh.invert_normals) {
        nx = e2[1] * e1[2] - e2[2] * e1[1];
        ny = e2[2] * e1[0] - e2[0] * e1[2];
        nz = e2[0] * e1[1] - e2[1] * e1[0];
      } else {
        nx = cx;
        ny = cy;
        nz = cz;
      }
      float l2 = nx * nx + ny * ny + nz * nz;
      if (l2 > 0.0f) {
        double inv = 1.0 / (double)std::sqrt(l2);
        nx = (float)(nx * inv);
        ny = (float)(ny * inv);
        nz = (float)(nz * inv);
      }
      DevLightTri T;
      T.cdf = acc;
      T.n[0] = nx;
      T.n[1] = ny;
      T.n[2] = nz;
      for (int c = 0; c < 3; ++c) {
        T.v0[c] = v[0][c];
        T.v1[c] = v[1][c];
        T.v2[c] = v[2][c];
        for (int k = 0; k < 3; ++k) {
          fmn[c] = (v[k][c] < fmn[c]) ? v[k][c] : fmn[c];
          fmx[c] = (fmx[c] < v[k][c]) ? v[k][c] : fmx[c];
        }
      }
      d->ltris.push_back(T);
    }
    L.area = L.tri_count > 0 ? acc : 0.0;
    if (L.area > 0.0) {
      double inv_area = 1.0 / L.area;
      for (size_t t = first; t < d->ltris.size(); ++t) d->ltris[t].cdf *= inv_area;
    }
    for (int c = 0; c < 3; ++c) {
      L.bmin[c] = fmn[c];
      L.bmax[c] = fmx[c];
    }
    double x = std::fabs(L.bmax[0] - L.bmin[0]), y = std::fabs(L.bmax[1] - L.bmin[1]), z = std::fabs(L.bmax[2] - L.bmin[2]);
    double bbox_area = 2.0 * (x * y + x * z + y * z);
    double lum = 0.0;
    for (int c = 0; c < 3; ++c) lum += kLumW[c] * ((kPi * L.area) * L.radiance[c]);
    L.intensity = lum / (kPi * bbox_area);
    d->lights.push_back(L);
    for (int c = 0; c < 3; ++c) d->rad.push_back(L.radiance[c]);
  }
  SKW_HIP(upload(&d->d_mats, d->mats), "upload materials");
  SKW_HIP(upload(&d->d_lights, d->lights), "upload lights");
  SKW_HIP(upload(&d->d_ltris, d->ltris), "upload light triangles");
  SKW_HIP(upload(&d->d_rad, d->rad), "upload light radiance");
  *out = d.release();
  return SKW_OK;
}

skw_status skw_dl_scene_destroy(skw_dl_scene *dl) {
  delete dl;
  return SKW_OK;
}

}  // extern "C"

namespace skw {
// Per-sample Le + direct lighting (3 doubles per sample, pixel * spp + pass) of a plain frame,
// before the renderer's sample rule; rays_out: primary, shadow, BSDF rays.
skw_status dl_sample_radiance(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                              const skw_dl_params *prm, double *d_sample_L, uint64_t *rays_out, hipStream_t st) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  FrameOut out;
  out.sample_L = d_sample_L;
  return run_frame(scene, dl, ParamPlan(), cam, prm, out, rays_out, st);
}
}  // namespace skw

extern "C" {

skw_status skw_dl_render(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                         const skw_dl_params *prm, float *d_rgb, uint64_t *rays_out, void *stream) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  if (!d_rgb) return fail(SKW_INVALID_ARG, "d_rgb must not be NULL");
  FrameOut out;
  out.rgb = d_rgb;
  return run_frame(scene, dl, ParamPlan(), cam, prm, out, rays_out, reinterpret_cast<hipStream_t>(stream));
}

skw_status skw_dl_render_profiled(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                                  const skw_dl_params *prm, float *d_rgb, int count_tests, skw_dl_profile *profile,
                                  void *stream) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  if (!d_rgb || !profile) return fail(SKW_INVALID_ARG, "d_rgb and profile must not be NULL");
  FrameOut out;
  out.rgb = d_rgb;
  return run_frame(scene, dl, ParamPlan(), cam, prm, out, nullptr, reinterpret_cast<hipStream_t>(stream), profile,
                   count_tests != 0);
}

skw_status skw_dl_param_count(const skw_dl_scene *dl, uint32_t *n) {
  if (!dl || !n) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  *n = 2 * dl->n_groups + (uint32_t)dl->lights.size();
  return SKW_OK;
}

skw_status skw_dl_render_replay(const skw_scene *scene, const skw_dl_scene *dl, const double *theta_eval,
                                const double *theta_sample, const skw_camera *cam, const skw_dl_params *prm,
                                double *d_image, uint64_t *rays_out, void *stream) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  if (!d_image) return fail(SKW_INVALID_ARG, "d_image must not be NULL");
  if ((s = check_theta(dl, theta_eval)) != SKW_OK || (s = check_theta(dl, theta_sample)) != SKW_OK) return s;
  ParamPlan pp;
  pp.active = true;
  pp.theta_e = theta_eval;
  pp.theta_s = theta_sample;
  const size_t n = 2 * dl->n_groups + dl->lights.size();
  pp.alias = n == 0 || std::memcmp(theta_eval, theta_sample, n * sizeof(double)) == 0;
  FrameOut out;
  out.img = d_image;
  return run_frame(scene, dl, pp, cam, prm, out, rays_out, reinterpret_cast<hipStream_t>(stream));
}

skw_status skw_dl_adjoint(const skw_scene *scene, const skw_dl_scene *dl, const double *theta,
                          const skw_camera *cam, const skw_dl_params *prm, const double *d_adjoint,
                          double *d_grad, uint64_t *rays_out, void *stream) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  if (!d_adjoint || !d_grad) return fail(SKW_INVALID_ARG, "d_adjoint and d_grad must not be NULL");
  if ((s = check_theta(dl, theta)) != SKW_OK) return s;
  ParamPlan pp;
  pp.active = true;
  pp.alias = true;
  pp.theta_e = theta;
  pp.theta_s = theta;
  FrameOut out;
  out.adj = d_adjoint;
  out.grad = d_grad;
  return run_frame(scene, dl, pp, cam, prm, out, rays_out, reinterpret_cast<hipStream_t>(stream));
}

skw_status skw_dl_render_tape(const skw_scene *scene, const skw_dl_scene *dl, const double *theta,
                              const skw_camera *cam, const skw_dl_params *prm, double *d_image,
                              uint64_t max_tape_bytes, skw_dl_tape **tape, uint64_t *rays_out, void *stream) {
  skw_status s = check_frame_args(scene, dl, cam, prm);
  if (s != SKW_OK) return s;
  if (!d_image || !tape) return fail(SKW_INVALID_ARG, "d_image and tape must not be NULL");
  if ((s = check_theta(dl, theta)) != SKW_OK) return s;
  const uint32_t n_lights = (uint32_t)dl->lights.size();
  const uint64_t batch = (uint64_t)prm->width * prm->height;
  skw_dl_tape *old = *tape;
  if (old && (old->n_groups != dl->n_groups || old->n_lights != n_lights || old->spp != prm->spp ||
              old->width != prm->width || old->height != prm->height))
    return fail(SKW_INVALID_ARG, "the tape to reuse has another frame shape");
  std::unique_ptr<skw_dl_tape> fresh;
  skw_dl_tape *t = old;
  if (!t) {
    fresh.reset(new skw_dl_tape());
    t = fresh.get();
    t->n_groups = dl->n_groups;
    t->n_lights = n_lights;
    t->stride = 6 + 3 * n_lights;
    t->spp = prm->spp;
    t->width = prm->width;
    t->height = prm->height;
    t->batch = batch;
    t->n_samples = batch * prm->spp;
    const uint64_t bytes = t->n_samples * (t->stride * sizeof(double) + sizeof(uint32_t));
    if (bytes > max_tape_bytes)
      return fail(SKW_OUT_OF_MEMORY, "the tape would exceed max_tape_bytes; use skw_dl_adjoint");
    hipError_t e = hipMalloc((void **)&t->d_j, t->n_samples * t->stride * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&t->d_group, t->n_samples * sizeof(uint32_t));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(SKW_OUT_OF_MEMORY, "no device memory for the tape; use skw_dl_adjoint");
    }
  }
  ParamPlan pp;
  pp.active = true;
  pp.alias = true;
  pp.theta_e = theta;
  pp.theta_s = theta;
  FrameOut out;
  out.img = d_image;
  out.tape = t;
  s = run_frame(scene, dl, pp, cam, prm, out, rays_out, reinterpret_cast<hipStream_t>(stream));
  if (s != SKW_OK) return s;
  if (fresh) *tape = fresh.release();
  return SKW_OK;
}

skw_status skw_dl_tape_adjoint(const skw_dl_tape *t, const double *d_adjoint, double *d_grad, void *stream) {
  if (!t || !d_adjoint || !d_grad) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SKW_HIP(launch_dl_clear_grad(d_grad, 2 * t->n_groups + t->n_lights, st), "gradient clear");
  DlTapeArgs a;
  a.tape_j = t->d_j;
  a.tape_group = t->d_group;
  a.stride = t->stride;
  a.n_groups = t->n_groups;
  a.n_lights = t->n_lights;
  a.spp = t->spp;
  a.width = t->width;
  a.height = t->height;
  a.batch = t->batch;
  a.n_samples = t->n_samples;
  a.adj = d_adjoint;
  a.grad = d_grad;
  SKW_HIP(launch_dl_tape_adjoint(a, st), "tape contraction");
  return SKW_OK;
}

skw_status skw_dl_tape_destroy(skw_dl_tape *t) {
  if (t) (void)hipDeviceSynchronize();
  delete t;
  return SKW_OK;
}

}  // extern "C"
