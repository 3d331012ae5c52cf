/* ORACLE — TEST INFRASTRUCTURE ONLY.
 * C interface of the CPU restatement of Skwarka's ray/photon hot path. Only the tests, smoke()
 * and the CPU-baseline leg of bench.py load it (through ctypes); the product library never does.
 * Each function names the reference code it restates in its source file. */
#ifndef ORC_API_H
#define ORC_API_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_accel orc_accel;
typedef struct orc_kdtree orc_kdtree;

/* Closest-hit record: ids after mesh cleaning, t from traversal, b1/b2 in mesh space. */
typedef struct {
  uint32_t prim, tri;
  double t, b1, b2;
} orc_hit;

/* World-space hit geometry plus the mesh-space values GetNextMinT needs. */
typedef struct {
  double p[3], ng[3], ns[3], cross[3], dot;
} orc_hit_geom;

orc_accel *orc_accel_create(uint32_t n_meshes, const uint64_t *mesh_vert_off, const float *verts,
                            const float *normals, const uint64_t *mesh_tri_off, const uint32_t *tris,
                            const uint8_t *mesh_flags, uint32_t n_prims, const uint32_t *prim_mesh,
                            const double *prim_xf);
void orc_accel_destroy(orc_accel *h);
void orc_accel_info(const orc_accel *h, uint64_t *out5);
void orc_accel_bounds(const orc_accel *h, double *out6);
uint64_t orc_accel_dump(const orc_accel *h, float *bbox, uint64_t *meta, uint32_t *tri_prim, uint32_t *tri_tri);
void orc_closest_hit(const orc_accel *h, uint64_t n, const double *rays, orc_hit *hits, orc_hit_geom *geom,
                     uint32_t *pops, uint32_t *tests, int nthreads);
void orc_any_hit(const orc_accel *h, uint64_t n, const double *rays, uint8_t *out, int nthreads);
void orc_brute_closest(const orc_accel *h, uint64_t n, const double *rays, orc_hit *hits, int nthreads);
double orc_next_min_t(const orc_accel *h, uint32_t prim, const orc_hit_geom *g, const double *dir3);
int orc_accel_intersect_one(const orc_accel *h, const double *ray, orc_hit *hit, orc_hit_geom *geom);
uint32_t orc_accel_prim_count(const orc_accel *h);
uint32_t orc_accel_mesh_tri_count(const orc_accel *h, uint32_t prim);
int orc_accel_mesh_inverted(const orc_accel *h, uint32_t prim);
void orc_accel_mesh_tri(const orc_accel *h, uint32_t prim, uint32_t tri, float *v9);

/* Direct-lighting render (DirectLightingIntegrator with matte / plastic materials and diffuse
 * area lights). mat_type: 0 matte (Lambertian), 1 plastic (Lambertian + Blinn microfacet with
 * FresnelDielectric(1.5, 1)); kd, ks 3 per prim; roughness per prim. Lights: prim index (its mesh
 * in world space) and RGB radiance. */
typedef struct orc_dl_scene orc_dl_scene;
orc_dl_scene *orc_dl_scene_create(const orc_accel *accel, const uint8_t *mat_type, const double *kd, const double *ks,
                                  const double *roughness, uint32_t n_lights, const uint32_t *light_prim,
                                  const double *light_radiance);
void orc_dl_scene_destroy(orc_dl_scene *h);
/* cam: origin[3], camera-to-world rotation[9] (row-major), x_tan, y_tan. Pixels [p0, p1);
 * rgb 3 floats per pixel of the range; counters: primary, shadow, BSDF rays cast. */
void orc_dl_render(const orc_accel *accel, const orc_dl_scene *dls, const double *cam, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples, uint64_t seed, uint64_t p0,
                   uint64_t p1, float *rgb, uint64_t *counters, uint32_t *primary_prim, uint32_t *primary_tri,
                   int nthreads);
/* Parameter groups: prim_group[p] in [0, n_groups) or 0xFFFFFFFF; theta = [reflectance G,
 * roughness G, emission scale per light]. */
orc_dl_scene *orc_dl_scene_create_groups(const orc_accel *accel, const uint8_t *mat_type, const double *kd,
                                         const double *ks, const double *roughness, uint32_t n_lights,
                                         const uint32_t *light_prim, const double *light_radiance, uint32_t n_groups,
                                         const uint32_t *prim_group);
uint32_t orc_dl_param_count(const orc_dl_scene *h);
/* Replayed estimator F(theta_eval; theta_sample): sampling decisions, pdfs, MIS weights and the
 * light CDF at theta_sample, BSDF values and emitted radiance at theta_eval. img: 3 doubles per
 * pixel of [p0, p1). */
void orc_dl_render_replay(const orc_accel *accel, const orc_dl_scene *dls, const double *theta_eval,
                          const double *theta_sample, const double *cam, uint32_t width, uint32_t height, uint32_t spp,
                          uint32_t n_light_samples, uint32_t n_bsdf_samples, uint64_t seed, uint64_t p0, uint64_t p1,
                          double *img, int nthreads);
/* grad[j] = sum over pixels of [p0, p1) and channels of adj * dF/dtheta_j at theta (forward-mode
 * dual numbers); img optional. */
void orc_dl_gradient(const orc_accel *accel, const orc_dl_scene *dls, const double *theta, const double *cam,
                     uint32_t width, uint32_t height, uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples,
                     uint64_t seed, uint64_t p0, uint64_t p1, const double *adj, double *grad, double *img,
                     int nthreads);

/* Single-function hooks for the oracle's own checks: BSDF pieces in the local frame (normal = z),
 * FresnelDielectric(1.5, 1), the lights CDF and area-light sampling of a DL scene. */
double orc_blinn_d(double e, const double *h);
double orc_fresnel_dielectric(double cos_i);
int orc_microfacet_eval(double e, const double *wo, const double *wi, double *scalar, double *fres);
double orc_microfacet_pdf(double e, const double *wo, const double *wi);
void orc_bsdf_local_eval(int plastic, const double *kd, const double *ks, double e, const double *wo,
                         const double *wi, double *f, double *pdf);
double orc_bsdf_local_sample(int plastic, const double *kd, const double *ks, double e, const double *wo, double u0,
                             double u1, double uc, double *wi, double *f);
void orc_dl_lights_cdf(const orc_dl_scene *h, const double *p, const double *n, double *cdf);
uint32_t orc_dl_light_sample(const orc_dl_scene *h, uint32_t l, const double *p, double tri_u, double u0, double u1,
                             double *wi, double *dist, double *pdf, int *lit);
double orc_dl_light_pdf(const orc_dl_scene *h, uint32_t l, uint32_t tri, const double *wi, double dist);
void orc_dl_light_info(const orc_dl_scene *h, uint32_t l, double *area, double *intensity);

typedef struct orc_irr_map orc_irr_map;
typedef struct orc_photon_map orc_photon_map;
/* PhotonLTEIntegrator::_FinalGather (PhotonLTEIntegrator.cpp:387-577) at the camera hit of every sample
 * of pixels [p0, p1) (the direct-lighting camera stream): radiance = 3 doubles per sample, sample
 * (pix - p0) * spp + s. gather_samples is rounded up to a square. indirect may be NULL. */
void orc_fg_render(const orc_accel *accel, const orc_dl_scene *dls, const orc_irr_map *irr,
                   const orc_photon_map *indirect, double scene_area, const double *cam, uint32_t width,
                   uint32_t height, uint32_t spp, uint32_t gather_samples, uint64_t seed, uint64_t p0, uint64_t p1,
                   double *radiance, uint64_t *gather_rays, int nthreads);

/* PhotonLTEIntegrator::_SurfaceRadiance for matte / plastic scenes: ((Le + direct) + caustic (zero:
 * no specular surfaces)) + final gather per sample, the renderer's sample rule on the sum, pixel
 * means as float rgb over pixels [p0, p1). irr may be NULL (no final gather). counters: primary,
 * shadow, BSDF, gather rays. */
void orc_pm_render(const orc_accel *accel, const orc_dl_scene *dls, const orc_irr_map *irr,
                   const orc_photon_map *indirect, double scene_area, const double *cam, uint32_t width,
                   uint32_t height, uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples,
                   uint32_t gather_samples, uint64_t seed, uint64_t p0, uint64_t p1, float *rgb, uint64_t *counters,
                   int nthreads);

/* Generic point kd-tree (KDTree<Point3D_d> / KDTree<Point3D_f>) for the reference's own tests. */
orc_kdtree *orc_kdtree_create_d(uint64_t n, const double *pts);
orc_kdtree *orc_kdtree_create_f(uint64_t n, const float *pts);
void orc_kdtree_destroy(orc_kdtree *h);
/* Tree-order permutation: perm[i] = original index of the point at tree position i; nodes as
 * the reference packs them (split float, 32-bit flags/right-child word). */
void orc_kdtree_dump(const orc_kdtree *h, uint32_t *perm, float *split, uint32_t *word);
/* filter: 0 = none, 1 = the reference test's CustomPointFilter (odd (int)(x+y+z)). Outputs
 * original point ids and d^2 in the processor's final array order; returns per-query count. */
void orc_kdtree_knn(const orc_kdtree *h, uint64_t nq, const double *q, const double *radius, const uint32_t *k,
                    uint32_t kmax, int filter, uint32_t *ids, double *d2, uint32_t *count, int nthreads);
void orc_kdtree_nearest(const orc_kdtree *h, uint64_t nq, const double *q, const double *radius, int filter,
                        uint32_t *id, int nthreads);

/* Photon map: KDTree<Photon> with PhotonFilter and the density estimators. */
orc_photon_map *orc_photon_map_create(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                                      const uint16_t *normal);
void orc_photon_map_destroy(orc_photon_map *h);
void orc_photon_map_dump(const orc_photon_map *h, uint32_t *perm, float *split, uint32_t *word);
/* _LookupPhotonIrradiance: k nearest with PhotonFilter inside r; E_ext/E_int float RGB. */
void orc_photon_irradiance(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                           double r, uint64_t paths, float *e_ext, float *e_int, uint32_t *ids, double *d2,
                           uint32_t *count, int nthreads);
/* Per query of orc_photon_irradiance: nodes visited, photons inside the radius, set size. */
void orc_photon_lookup_counts(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                              double r, uint32_t *counts, int nthreads);
/* _LookupCausticRadiance with a Lambertian BRDF of albedo rho (reflection side only). */
void orc_photon_caustic(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, const double *wo,
                        uint32_t k, double r, uint64_t paths, double rho, double *radiance, int nthreads);
/* Cone-filtered irradiance (Jensen: w = 1 - d/(kc*r_max), normalised by 1 - 2/(3 kc)). */
void orc_photon_cone(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                     double r, uint64_t paths, double kc, float *e_ext, int nthreads);

/* GetNearestPoint with PhotonFilter within r: original index or 0xFFFFFFFF. */
void orc_photon_nearest(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, double r, uint32_t *id,
                        int nthreads);
/* The map's photons in tree order (GetAllPoints): positions, dir and normal codes (either may be NULL). */
void orc_photon_map_points(const orc_photon_map *h, float *pos, uint16_t *dir, uint16_t *normal);
uint64_t orc_photon_map_size(const orc_photon_map *h);

/* Irradiance photon map (_ConstructIrradiancePhotonMap, PhotonLTEIntegrator.cpp:171-245): every 10th
 * indirect photon (tree order) with E_ext / E_int = caustic + direct + indirect disc estimates (k = 100,
 * radii from the scene area), a kd-tree over them, and the lookup distance covering 99.9 % of the
 * indirect photons. caustic may be NULL. */
orc_irr_map *orc_irr_map_create(const orc_photon_map *caustic, const orc_photon_map *direct,
                                const orc_photon_map *indirect, uint64_t caustic_paths, uint64_t direct_paths,
                                uint64_t indirect_paths, double scene_area, int nthreads);
orc_irr_map *orc_irr_map_create_from(uint64_t n, const float *pos, const uint16_t *normal, const float *e_ext,
                                     const float *e_int, double max_dist);
void orc_irr_map_destroy(orc_irr_map *m);
void orc_irr_map_info(const orc_irr_map *m, uint64_t *n, double *max_dist);
/* Creation order: positions (3 floats), normal codes, E_ext and E_int (3 floats each). */
void orc_irr_map_export(const orc_irr_map *m, float *pos, uint16_t *normal, float *e_ext, float *e_int);
void orc_irr_map_nearest(const orc_irr_map *m, uint64_t nq, const double *p, const double *n, double r, uint32_t *id,
                         int nthreads);

/* CompressedDirection codec (Source/RayLibs/Math/CompressedDirection.{h,cpp}). */
uint16_t orc_cdir_encode(double x, double y, double z);

/* Photon shooting (PhotonLTEIntegrator::ShootPhotons with the shooting pipeline): paths in rounds
 * of paths_per_round, maps 0 caustic / 1 direct / 2 indirect, max_photons[3]. */
typedef struct orc_photon_set orc_photon_set;
orc_photon_set *orc_photon_shoot(const orc_accel *accel, const orc_dl_scene *dls, uint64_t max_paths,
                                 const uint64_t *max_photons, uint32_t paths_per_round, uint64_t seed, int nthreads);
void orc_photon_set_info(const orc_photon_set *s, uint64_t *counts, uint64_t *paths, uint64_t *total_paths);
void orc_photon_set_export(const orc_photon_set *s, uint32_t which, float *pos, float *weight, uint16_t *dir,
                           uint16_t *normal);
void orc_photon_set_destroy(orc_photon_set *s);
double orc_radical_inverse(uint32_t n, uint32_t base);
void orc_cdir_table(double *out_65536x3);

/* Raw 32-bit outputs of std::mt19937 (boost::mt19937 is the same engine). */
void orc_mt19937(uint32_t seed, uint64_t n, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
