// skw_pm_render: the photon-mapping integrator's per-sample radiance composed on the device,
// PhotonLTEIntegrator::_SurfaceRadiance (Raytracer/LTEIntegrators/PhotonLTEIntegrator/
// PhotonLTEIntegrator.cpp:320-385) = ((Le + ComputeDirectLighting) + _LookupCausticRadiance) +
// _FinalGather at every camera hit, then the renderer's sample rule on the sum and the pixel mean.
// The direct-lighting frame and the final-gather frame run as they do alone (each writes its
// per-sample values instead of a film), and one kernel composes them.
#include <hip/hip_runtime.h>

#include <cstring>

#include "internal.hpp"

using namespace skw;

struct skw_dl_scene;

namespace skw {
skw_status dl_sample_radiance(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                              const skw_dl_params *prm, double *d_sample_L, uint64_t *rays_out, hipStream_t st);
}  // namespace skw

extern "C" {

skw_status skw_pm_render(const skw_scene *scene, const skw_dl_scene *dl, const skw_irradiance_map *irr,
                         const skw_photon_map *caustic, const skw_photon_map *indirect, double scene_area,
                         const skw_camera *cam, const skw_pm_params *prm, float *d_rgb, uint64_t *rays_out,
                         void *stream) {
  if (!scene || !dl || !cam || !prm || !d_rgb) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (caustic && caustic->n > 0)
    return fail(SKW_UNSUPPORTED,
                "caustic photons need specular surfaces; with matte and plastic materials the caustic map is empty");
  if (prm->width == 0 || prm->height == 0 || prm->spp == 0) return fail(SKW_INVALID_ARG, "empty image");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const uint64_t n = (uint64_t)prm->width * prm->height * prm->spp;
  skw_dl_params dp;
  std::memset(&dp, 0, sizeof(dp));
  dp.width = prm->width;
  dp.height = prm->height;
  dp.spp = prm->spp;
  dp.light_samples = prm->light_samples;
  dp.bsdf_samples = prm->bsdf_samples;
  dp.seed = prm->seed;
  double *s_dl = nullptr, *s_fg = nullptr;
  SKW_HIP(pool_alloc((void **)&s_dl, n * 3 * sizeof(double), st), "per-sample buffers");
  if (irr) {
    const hipError_t e = pool_alloc((void **)&s_fg, n * 3 * sizeof(double), st);
    if (e != hipSuccess) {
      pool_free(s_dl, st);
      return hip_fail(e, "per-sample buffers");
    }
  }
  uint64_t rays[3] = {0, 0, 0}, grays[2] = {0, 0};
  skw_status rs = dl_sample_radiance(scene, dl, cam, &dp, s_dl, rays_out ? rays : nullptr, st);
  if (rs == SKW_OK && irr) {
    skw_fg_params fp;
    std::memset(&fp, 0, sizeof(fp));
    fp.width = prm->width;
    fp.height = prm->height;
    fp.spp = prm->spp;
    fp.gather_samples = prm->gather_samples;
    fp.seed = prm->seed;
    rs = skw_dl_final_gather(scene, dl, irr, indirect, scene_area, cam, &fp, nullptr, s_fg, rays_out ? grays : nullptr,
                             stream);
  }
  hipError_t e = hipSuccess;
  if (rs == SKW_OK) e = launch_pm_combine(s_dl, s_fg, (uint64_t)prm->width * prm->height, prm->spp, d_rgb, st);
  pool_free(s_dl, st);
  pool_free(s_fg, st);
  if (rs != SKW_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "photon-mapping composition");
  if (rays_out) {
    rays_out[0] = rays[0];
    rays_out[1] = rays[1];
    rays_out[2] = rays[2];
    rays_out[3] = grays[1];
  }
  return SKW_OK;
}

}  // extern "C"
