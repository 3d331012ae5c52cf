/*
 * skwarka_rt — MI355X (gfx950) closest-hit / occlusion traversal of Skwarka's ternary triangle
 * tree and the photon-map kd-tree gather, behind a plain C ABI.
 *
 * Every entry point returns an skw_status; skw_last_error() gives the thread-local message of the
 * last failure. Handles are immutable after creation; calls on different streams may overlap.
 * Device pointers ("d_" prefix) must be HIP device (or managed) memory; "stream" is a hipStream_t
 * (NULL = the null stream). Work is stream-ordered: outputs are ready when the stream is.
 *
 * Reference interfaces replaced (paths relative to the Skwarka source tree):
 *   Source/RayLibs/Raytracer/Core/TriangleAccelerator.h:41-62   TriangleAccelerator ctor / Intersect /
 *                                                              IntersectTest / GetWorldBounds
 *   Source/RayLibs/Raytracer/Core/Scene.h:77,85,140-148         Scene::Intersect / Scene::IntersectTest
 *   Source/RayLibs/Raytracer/Core/KDTree.h:45-119               KDTree<T> ctor / GetNearestPoints /
 *                                                              GetNearestPoint
 *   Source/RayLibs/Raytracer/LTEIntegrators/PhotonLTEIntegrator/PhotonLTEIntegrator.cpp:127-169, 579-636
 *                                                              _LookupPhotonIrradiance / _LookupCausticRadiance
 *   Source/RayLibs/Raytracer/Core/CoreUtils.h:139-149           CoreUtils::GetNextMinT (hit record fields)
 */
#ifndef SKWARKA_RT_H
#define SKWARKA_RT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SKW_OK = 0,
  SKW_INVALID_ARG = 1,
  SKW_OUT_OF_MEMORY = 2,
  SKW_HIP_ERROR = 3,
  SKW_UNSUPPORTED = 4,
  SKW_NO_DEVICE = 5
} skw_status;

/* Ray as the reference's Ray (Source/RayLibs/Math/Ray.h:31-60): origin, direction (normalised by
 * the caller), [tmin, tmax]. 64 bytes. */
typedef struct {
  double o[3];
  double d[3];
  double tmin, tmax;
} skw_ray;

/* Closest hit: primitive index (scene order), triangle index inside the primitive's mesh (after
 * the mesh drops triangles with out-of-range or repeated vertex indices), ray parameter t, and the
 * barycentrics b1, b2 of vertices 1 and 2 computed in mesh space. A miss has
 * prim = tri = 0xFFFFFFFF, t = +inf, b1 = b2 = 0. 32 bytes. */
typedef struct {
  uint32_t prim, tri;
  double t, b1, b2;
} skw_hit;

/* Hit geometry (TriangleAccelerator.cpp:322-345): world point, geometric and shading normals, and
 * the mesh-space cross/dot values GetNextMinT uses to keep a spawned ray off its own triangle. */
typedef struct {
  double p[3], ng[3], ns[3], cross[3], dot;
} skw_hit_geom;

/* Scene description. Meshes hold float vertices and uint32 vertex-index triples; a primitive
 * places a mesh in the world through a mesh_to_world matrix and its inverse (row-major 4x4
 * doubles, 32 per primitive; NULL = identity for all). A mesh shared by several primitives is
 * instanced (one level), as in TriangleAccelerator.cpp:54-102. */
#define SKW_MESH_INVERT_NORMALS 1u /* TriangleMesh m_invert_normals */
#define SKW_MESH_SHADING_NORMALS 2u /* TriangleMesh m_use_shading_normals: interpolate `normals` */
typedef struct {
  uint32_t n_meshes;
  const uint64_t *mesh_vert_offset; /* n_meshes + 1 */
  const float *verts;               /* 3 per vertex */
  const float *normals;             /* 3 per vertex shading normals, or NULL */
  const uint64_t *mesh_tri_offset;  /* n_meshes + 1 */
  const uint32_t *tris;             /* 3 per triangle, mesh-local vertex indices */
  const uint8_t *mesh_flags;        /* per mesh SKW_MESH_* bits, or NULL (all 0) */
  uint32_t n_prims;
  const uint32_t *prim_mesh;        /* n_prims */
  const double *prim_xform;         /* 32 per prim (matrix, inverse), or NULL */
} skw_scene_desc;

typedef struct skw_scene skw_scene;
typedef struct skw_photon_map skw_photon_map;

/* Scene flags. */
#define SKW_SCENE_HOST_ONLY 1u /* build the tree on the host only (no device upload) */
#define SKW_SCENE_DEVICE_BUILD 2u /* build the tree on the current HIP device */

/* TriangleAccelerator::TriangleAccelerator (TriangleAccelerator.cpp:22-147, Node::Node and
 * _DetermineBestSplit :505-729): builds the ternary tree (same split search, partition and limits)
 * and uploads the flattened arrays to the current HIP device. By default the tree is built on the
 * host; with SKW_SCENE_DEVICE_BUILD it is built on the device, level by level, node for node the
 * host's tree (scenes with instanced meshes or non-finite vertices: SKW_UNSUPPORTED). */
skw_status skw_scene_create(const skw_scene_desc *desc, uint32_t flags, skw_scene **out);
skw_status skw_scene_destroy(skw_scene *scene);

/* TriangleAccelerator::GetWorldBounds (TriangleAccelerator.cpp:149-152): min xyz, max xyz. */
skw_status skw_scene_bounds(const skw_scene *scene, double out_minmax[6]);

/* Counts: [0] nodes, [1] triangles, [2] max depth, [3] instances, [4] leaves, [5] stack depth the
 * traversal kernels reserve. */
skw_status skw_scene_info(const skw_scene *scene, uint64_t out[6]);

/* Host copy of the flattened tree in depth-first pre-order (children left, middle, right): per
 * node bbox (6 floats) and meta (flags, child mask, tri begin, tri end, inst begin, inst end);
 * per triangle position its (prim, tri) ids (prim = 0xFFFFFFFF for instanced meshes). Any output
 * may be NULL. Returns the node count through n_nodes. */
skw_status skw_scene_export_tree(const skw_scene *scene, float *bbox, uint64_t *meta, uint32_t *tri_prim,
                                 uint32_t *tri_tri, uint64_t *n_nodes);

/* Scene::Intersect / TriangleAccelerator::Intersect + _NodeIntersect (TriangleAccelerator.cpp:154-352),
 * batched: one hit record per ray. */
skw_status skw_closest_hit(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits,
                           void *stream);

/* skw_closest_hit that also records, per ray, the node boxes tested (d_pops) and the triangles
 * tested (d_tests), in the traversal order of _NodeIntersect; used for the roofline's algorithmic
 * byte count and to check visit order. Either counter array may be NULL. */
skw_status skw_closest_hit_stats(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits,
                                 uint32_t *d_pops, uint32_t *d_tests, void *stream);

/* skw_closest_hit that allocates nothing: no counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same kernels and bit-identical hit records as skw_closest_hit. Its
 * working memory (the ray-claim counter and, for trees deeper than 128 levels, the traversal
 * stack) is a fixed 16 MiB device array per GPU, held until the process ends; a tree whose
 * stack does not fit it (more than 65,535 entries per ray, beyond any tree the builder makes) is
 * rejected with SKW_UNSUPPORTED. Every call on one GPU shares that array: calls that overlap on one
 * GPU (on different streams, or from different threads without ordering) race and must not be
 * issued. */
skw_status skw_closest_hit_noalloc(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, skw_hit *d_hits,
                                   void *stream);

/* Scene::IntersectTest / _NodeIntersectTest (TriangleAccelerator.cpp:354-468): 1 if occluded. */
skw_status skw_any_hit(const skw_scene *scene, const skw_ray *d_rays, uint64_t n, uint8_t *d_occluded,
                       void *stream);

/* Hit geometry for closest-hit records (TriangleAccelerator.cpp:309-345, TriangleMesh.cpp:439-500).
 * Misses get zeros. */
skw_status skw_hit_geometry(const skw_scene *scene, const skw_ray *d_rays, const skw_hit *d_hits, uint64_t n,
                            skw_hit_geom *d_geom, void *stream);

/* Photon map: KDTree<Photon> (KDTree.h:391-473; Photon layout PhotonInternalTypes.h:35-59).
 * dir / normal are CompressedDirection ids. Built on the host (libstdc++ nth_element median split),
 * uploaded in tree order; results refer to the caller's photon indices. */
skw_status skw_photon_map_create(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                                 const uint16_t *normal, uint32_t flags, skw_photon_map **out);
skw_status skw_photon_map_destroy(skw_photon_map *map);
/* Photons in the map and the depth of its deepest node (host outputs, either may be NULL). */
skw_status skw_photon_map_info(const skw_photon_map *map, uint64_t *n, uint32_t *max_depth);
/* KDTree<Photon>::KDTree (KDTree.h:391-473) built on the device from device arrays: the same
 * tree, node for node, as skw_photon_map_create (both follow a libstdc++ build of the reference:
 * std::nth_element's introselect, restated exactly). d_weight, d_dir, d_normal may be NULL (0). */
skw_status skw_photon_map_build(uint64_t n, const float *d_pos, const float *d_weight, const uint16_t *d_dir,
                                const uint16_t *d_normal, skw_photon_map **out, void *stream);

/* Tree-order permutation and packed nodes (split float, flags/right-child word) on the host. */
skw_status skw_photon_map_export(const skw_photon_map *map, uint32_t *perm, float *split, uint32_t *word);

/* PhotonLTEIntegrator::_LookupPhotonIrradiance (PhotonLTEIntegrator.cpp:127-169): for each query point
 * and unit normal (3 doubles each), up to k nearest photons inside radius r that pass PhotonFilter
 * (PhotonInternalTypes.h:113-142), then the disc estimate split by side. Optional d_ids / d_d2
 * (k per query) receive the photon set in the bounded max-heap's final order, d_count the set size.
 * d_radiance (optional, 3 floats) gets rho/pi * E_ext for a Lambertian surface of albedo rho. */
skw_status skw_photon_irradiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                 uint64_t nq, uint32_t k, double r, uint64_t paths, double rho, float *d_e_ext,
                                 float *d_e_int, float *d_radiance, uint32_t *d_ids, double *d_d2,
                                 uint32_t *d_count, void *stream);

/* skw_photon_irradiance's lookups and outputs (no set returned) with the work of each lookup
 * counted: d_stats gets 3 uint32 per query = nodes visited (KDTree.h:545-575), photons inside the
 * radius that reach the PhotonFilter (KDTree.h:549-551, PhotonInternalTypes.h:128-134) and the
 * final set size. A measurement variant: the same walk as the plain call, slower. */
skw_status skw_photon_irradiance_stats(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                       uint64_t nq, uint32_t k, double r, uint64_t paths, double rho, float *d_e_ext,
                                       float *d_e_int, float *d_radiance, uint32_t *d_stats, void *stream);

/* PhotonLTEIntegrator::_LookupCausticRadiance + _PhotonKernel (PhotonLTEIntegrator.cpp:579-636) for a
 * Lambertian surface of albedo rho seen from direction d_wo (3 doubles per query, unit): the
 * PhotonFilter'd k-nearest set weighted by Simpson's kernel with the running maximum of d^2 in the
 * set's array order, / (paths * max d^2) with the n/(n-0.5) rule. d_radiance: 3 doubles per query. */
skw_status skw_photon_caustic_radiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                       const double *d_wo, uint64_t nq, uint32_t k, double r, uint64_t paths,
                                       double rho, double *d_radiance, void *stream);

/* Cone-filtered irradiance over the same PhotonFilter'd k-nearest set (Jensen's cone filter,
 * weight 1 - d/(kc * r_max), normalised by 1 - 2/(3 kc); photons from the normal's side only).
 * No counterpart in the reference, whose estimators are the disc (_LookupPhotonIrradiance) and
 * the Simpson kernel (_LookupCausticRadiance). d_e_ext: 3 floats per query. */
skw_status skw_photon_cone_irradiance(const skw_photon_map *map, const double *d_points, const double *d_normals,
                                      uint64_t nq, uint32_t k, double r, uint64_t paths, double kc, float *d_e_ext,
                                      void *stream);

/* KDTree::GetNearestPoints with or without PhotonFilter (filter = 0 none, 1 PhotonFilter at the
 * query normal). d_normals may be NULL when filter = 0. */
skw_status skw_photon_knn(const skw_photon_map *map, const double *d_points, const double *d_normals, uint64_t nq,
                          uint32_t k, double r, int filter, uint32_t *d_ids, double *d_d2, uint32_t *d_count,
                          void *stream);

/* KDTree::GetNearestPoint (KDTree.h:484-503) with the same filter choice; 0xFFFFFFFF = none. */
skw_status skw_photon_nearest(const skw_photon_map *map, const double *d_points, const double *d_normals,
                              uint64_t nq, double r, int filter, uint32_t *d_id, void *stream);

/* ---------------------------------------------------------------- direct lighting
 * DirectLightingLTEIntegrator / DirectLightingIntegrator over a scene (Raytracer/Core/
 * DirectLightingIntegrator.cpp:54-366, LTEIntegrators/DirectLightingLTEIntegrator.cpp:33-60): camera ray,
 * emitted light of area-light surfaces, light sampling with IrradianceLightsSamplingStrategy and
 * occlusion rays, BSDF sampling with closest-hit rays, power-heuristic MIS. Materials are
 * MatteMaterial (sigma 0: Lambertian) and PlasticMaterial (Lambertian + Blinn microfacet with
 * FresnelDielectric(1.5, 1)); lights are DiffuseAreaLightSource on primitives whose meshes are
 * in world space. Runs as a wavefront of kernels; all rays use the scene's traversal kernels. */
typedef struct skw_dl_scene skw_dl_scene;

typedef struct {
  const uint8_t *mat_type;         /* per prim: 0 matte, 1 plastic */
  const double *kd;                /* 3 per prim, clamped to [0,1] */
  const double *ks;                /* 3 per prim, clamped to [0,1] (plastic) */
  const double *roughness;         /* per prim (plastic), clamped to [0.001, 1] */
  uint32_t n_lights;
  const uint32_t *light_prim;      /* prim index of each area light */
  const double *light_radiance;    /* 3 per light */
  /* Parameter groups for skw_dl_render_replay / skw_dl_adjoint (optional: n_groups 0 and
   * prim_group NULL). A primitive in group g takes gray reflectance theta[g] (as kd) and roughness
   * theta[G + g]; light l takes radiance theta[2G + l] * light_radiance[l]. */
  uint32_t n_groups;               /* G, at most 64 */
  const uint32_t *prim_group;      /* per prim: group index or SKW_NO_GROUP */
} skw_dl_desc;

#define SKW_NO_GROUP 0xFFFFFFFFu

/* Pinhole camera (PerspectiveCamera::GenerateRay with lens radius 0): origin, camera-to-world
 * rotation (row-major), tan of half the x field of view, and y tan = x tan * height / width. */
typedef struct {
  double origin[3];
  double rot[9];
  double x_tan, y_tan;
} skw_camera;

/* Sample stream: dimension d of sample s of pixel p is splitmix64-derived
 * u = (mix(seed ^ mix((p*spp + s)*64 + d)) >> 11) * 2^-53; dims 0-1 pixel jitter, 2.. light
 * choice, then light position pairs, then BSDF component, then BSDF direction pairs. */
typedef struct {
  uint32_t width, height, spp;
  uint32_t light_samples, bsdf_samples; /* per hit; at most 8 each */
  uint64_t seed;
} skw_dl_params;

skw_status skw_dl_scene_create(const skw_scene *scene, const skw_dl_desc *desc, skw_dl_scene **out);
skw_status skw_dl_scene_destroy(skw_dl_scene *dl);

/* Renders the image: d_rgb = width*height*3 floats (mean radiance per pixel over spp samples;
 * NaN, infinite or negative-luminance samples count as black, SamplerBasedRenderer.cpp:458-478).
 * rays_out (host, optional, 3 values): primary, occlusion and BSDF rays cast. Work is enqueued
 * on `stream`; the call returns after the frame has been enqueued and the ray counts read. */
skw_status skw_dl_render(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                         const skw_dl_params *params, float *d_rgb, uint64_t *rays_out, void *stream);

/* Per-frame profile of skw_dl_render's wavefront: HIP-event time of each stage summed over the
 * sample passes, the rays cast per class, (when counted) the node boxes and triangles the
 * traversal tested per class, and the frame's launches: each launch covers a group of sample
 * passes of every pixel and traces each ray class once. Classes: 0 primary, 1 occlusion, 2 BSDF. */
typedef struct {
  double stage_ms[6]; /* camera, primary trace, shade, occlusion trace, BSDF trace, resolve */
  uint64_t rays[3];
  uint64_t node_tests[3];
  uint64_t tri_tests[3];
  uint64_t launches; /* pass groups of the frame (spp / passes per launch) */
} skw_dl_profile;

/* skw_dl_render with a profile (no counterpart in the reference; for measurement). count_tests
 * selects the traversal's counting variant (same walk, slower): use a separate frame for timing. */
skw_status skw_dl_render_profiled(const skw_scene *scene, const skw_dl_scene *dl, const skw_camera *cam,
                                  const skw_dl_params *params, float *d_rgb, int count_tests, skw_dl_profile *profile,
                                  void *stream);

/* ---------------------------------------------------------------- differentiable direct lighting
 * No counterpart in the reference. Parameters theta (host, 2G + n_lights doubles): reflectance
 * per group, roughness per group, emission scale per light (see skw_dl_desc).
 *
 * The replayed estimator F(theta_eval; theta_sample) runs skw_dl_render's wavefront with every
 * sampling decision, pdf, MIS weight and light-choice CDF taken at theta_sample and the BSDF
 * values and emitted radiance taken at theta_eval, so F(theta; theta) is skw_dl_render at theta
 * and F(., theta_sample) is smooth in theta_eval for a fixed set of paths. Its gradient at
 * theta_eval = theta_sample is an unbiased estimate of the gradient of the expected image (the
 * MIS weights sum to one over the strategies; these parameters move no geometry). */

/* Number of parameters, 2G + n_lights. */
skw_status skw_dl_param_count(const skw_dl_scene *dl, uint32_t *n);

/* d_image = width*height*3 doubles: the replayed estimator's mean radiance per pixel. */
skw_status skw_dl_render_replay(const skw_scene *scene, const skw_dl_scene *dl, const double *theta_eval,
                                const double *theta_sample, const skw_camera *cam, const skw_dl_params *params,
                                double *d_image, uint64_t *rays_out, void *stream);

/* Adjoint of skw_dl_render_replay at theta_eval = theta_sample = theta: replays the same paths
 * (recomputing them, nothing is kept from a forward call) and writes
 *   d_grad[j] = sum over pixels and channels of d_adjoint[pixel, c] * dImage[pixel, c] / dtheta[j]
 * (d_adjoint = width*height*3 doubles on the device, e.g. dLoss/dImage; d_grad = 2G + n_lights
 * doubles on the device, overwritten). Accumulates with FP64 atomics (LDS per workgroup, then
 * global), so the summation order, and the last bits, vary from call to call. */
skw_status skw_dl_adjoint(const skw_scene *scene, const skw_dl_scene *dl, const double *theta,
                          const skw_camera *cam, const skw_dl_params *params, const double *d_adjoint,
                          double *d_grad, uint64_t *rays_out, void *stream);

/* Store-instead-of-recompute form of the adjoint. skw_dl_render_tape renders
 * skw_dl_render_replay(theta, theta) into d_image and keeps each sample's derivatives on the device
 * (a tape of (6 + 3 n_lights) doubles + 4 bytes per sample, e.g. 310 MB for 512x512 x 8 spp with
 * 4 lights); skw_dl_tape_adjoint then contracts the tape with any dLoss/dImage (one pass over the
 * tape, no rays), writing the same gradient skw_dl_adjoint computes. When the tape would exceed
 * max_tape_bytes or device memory, skw_dl_render_tape returns SKW_OUT_OF_MEMORY and the caller
 * recomputes with skw_dl_adjoint instead. *tape: NULL to allocate a new tape, or a tape of the same
 * frame shape to overwrite (no allocation). */
typedef struct skw_dl_tape skw_dl_tape;
skw_status skw_dl_render_tape(const skw_scene *scene, const skw_dl_scene *dl, const double *theta,
                              const skw_camera *cam, const skw_dl_params *params, double *d_image,
                              uint64_t max_tape_bytes, skw_dl_tape **tape, uint64_t *rays_out, void *stream);
skw_status skw_dl_tape_adjoint(const skw_dl_tape *tape, const double *d_adjoint, double *d_grad, void *stream);
skw_status skw_dl_tape_destroy(skw_dl_tape *tape);

/* ---------------------------------------------------------------- photon final gather
 * The photon integrator's indirect term: an irradiance photon map built from the direct, indirect
 * and caustic photon maps, and the final gather that reads it through gather rays. */

/* PhotonLTEIntegrator::_ConstructIrradiancePhotonMap (PhotonLTEIntegrator.cpp:171-245): every 10th
 * photon of the indirect map in tree order becomes an irradiance photon (position and normal of that
 * photon) whose external / internal irradiance is the sum, in float, of the caustic, direct and
 * indirect disc estimates at it (_LookupPhotonIrradiance with k = 100 and the radius holding 100
 * photons of each map on average over scene_area; IrradiancePhotonProcess, PhotonInternalTypes.h:
 * 269-288); the irradiance photons get a kd-tree, and the lookup distance is the 0.1 % percentile
 * of the distances from every 11th indirect photon to its nearest irradiance photon (infinite when
 * that distance is 0). The estimates and lookups run on the device; the tree is built on the host.
 * caustic may be NULL; direct and indirect are required (device maps). scene_area: the summed
 * areas of the scene's meshes (TriangleMesh::GetArea per primitive). */
typedef struct skw_irradiance_map skw_irradiance_map;
typedef struct {
  const skw_photon_map *caustic, *direct, *indirect;
  uint64_t caustic_paths, direct_paths, indirect_paths;
  double scene_area;
} skw_irradiance_map_desc;
skw_status skw_irradiance_map_create(const skw_irradiance_map_desc *desc, skw_irradiance_map **out, void *stream);
/* An irradiance photon map from given photons (KDTree<IrradiancePhoton> over them): positions (3
 * floats), normal ids (CompressedDirection), external and internal irradiance (3 floats each), and
 * the lookup distance. Host arrays. */
skw_status skw_irradiance_map_create_from(uint64_t n, const float *pos, const uint16_t *normal, const float *e_ext,
                                          const float *e_int, double max_lookup_dist, skw_irradiance_map **out);
skw_status skw_irradiance_map_destroy(skw_irradiance_map *map);
/* Irradiance photons and lookup distance (host outputs, any may be NULL; arrays in creation order). */
skw_status skw_irradiance_map_info(const skw_irradiance_map *map, uint64_t *n, double *max_lookup_dist);
skw_status skw_irradiance_map_export(const skw_irradiance_map *map, float *pos, uint16_t *normal, float *e_ext,
                                     float *e_int);

/* PhotonLTEIntegrator::_FinalGather (PhotonLTEIntegrator.cpp:387-577) at the camera hit of every
 * sample of a frame (the direct-lighting camera and sample stream, so the shading points are those
 * of skw_dl_render with the same width, height, spp and seed): the 32 nearest indirect photons
 * with PhotonFilter, gather_samples BSDF samples and gather_samples samples of 10-degree cones
 * around those photons' directions with power-heuristic weights, one closest-hit gather ray each,
 * and at each gather hit the nearest irradiance photon (IrradiancePhotonFilter, within the map's
 * lookup distance) times the hit surface's total reflection scattering / pi. gather_samples is
 * rounded up to a square as the reference does without a Sample; its stratified 1D / 2D
 * sequences are jittered from a counter-based stream of their own. indirect may be NULL (no
 * photon-direction samples). Outputs (device, any may be NULL): d_rgb = mean over spp per pixel
 * (3 floats), d_radiance = every sample's value (3 doubles, sample pixel * spp + s). rays_out
 * (host, optional, 2 values): camera rays, gather rays. */
typedef struct {
  uint32_t width, height, spp;
  uint32_t gather_samples;
  uint64_t seed;
} skw_fg_params;
skw_status skw_dl_final_gather(const skw_scene *scene, const skw_dl_scene *dl, const skw_irradiance_map *irr,
                               const skw_photon_map *indirect, double scene_area, const skw_camera *cam,
                               const skw_fg_params *params, float *d_rgb, double *d_radiance, uint64_t *rays_out,
                               void *stream);

/* The photon-mapping render: PhotonLTEIntegrator::_SurfaceRadiance (PhotonLTEIntegrator.cpp:320-385)
 * per camera sample = ((Le + ComputeDirectLighting) + _LookupCausticRadiance) + _FinalGather, then
 * the renderer's sample rule (NaN, negative luminance, infinite: dropped) on the sum and the pixel
 * mean into d_rgb (3 floats per pixel). The direct-lighting part is skw_dl_render's (its sample
 * stream, light_samples + bsdf_samples per hit), the final gather skw_dl_final_gather's over irr and
 * the indirect map (its own stream; irr NULL: no final gather, as the reference without an
 * irradiance map). Matte and plastic surfaces scatter no photon specularly, so the caustic map is
 * empty and its lookup adds a zero spectrum (PhotonLTEIntegrator.cpp:585-586): caustic must be NULL
 * or hold no photons (SKW_UNSUPPORTED otherwise). rays_out (optional, 4): camera, shadow, BSDF and
 * gather rays. */
typedef struct {
  uint32_t width, height, spp;
  uint32_t light_samples, bsdf_samples; /* per hit, 1..8 each */
  uint32_t gather_samples;              /* 1..676, rounded up to a square */
  uint64_t seed;
} skw_pm_params;
skw_status skw_pm_render(const skw_scene *scene, const skw_dl_scene *dl, const skw_irradiance_map *irr,
                         const skw_photon_map *caustic, const skw_photon_map *indirect, double scene_area,
                         const skw_camera *cam, const skw_pm_params *params, float *d_rgb, uint64_t *rays_out,
                         void *stream);

/* ---------------------------------------------------------------- photon shooting
 * PhotonLTEIntegrator::ShootPhotons (PhotonLTEIntegrator.cpp:247-288) with the shooting pipeline
 * (PhotonShootingPipeline.cpp:48-267) from the area lights of a direct-lighting scene: light
 * chosen by the lights' power CDF, DiffuseAreaLightSource::SamplePhoton, photons deposited at
 * every intersection (map 1 direct at the first, map 0 caustic while the path is specular, map 2
 * indirect otherwise), BSDF sampling and Russian roulette. Samples of path i: radical inverses of
 * i + 1 as the reference (bases 2, 3, 5, 7, 11, 13 and 17, 19, 23 at the first intersection); the
 * reference draws later bounces and the roulette from a per-chunk generator, here a counter-based
 * stream of (seed, path). Paths run in rounds of paths_per_round in path order; a map takes a
 * round's photons and counts its paths unless it already held max_photons[map] when the round
 * began (the pipeline's input filter decides per 4096-path chunk); shooting stops at max_paths,
 * when every map is full, or after 10^7 paths when the maps that are not full are all empty.
 * Photons of a map are in (path, intersection) order. max_photons are taken as given: pass the
 * integrator's clamped values, as the reference's constructor leaves them (PhotonLTEIntegrator.cpp:
 * 42-49: a limit of 0 becomes MAX_PHOTONS_IN_MAP = 40,000,000 and larger limits are clamped to it);
 * a limit of 0 here means the map takes no photons. */
typedef struct skw_photon_set skw_photon_set;
typedef struct {
  uint64_t max_paths;
  uint64_t max_photons[3]; /* caustic, direct, indirect */
  uint32_t paths_per_round;
  uint64_t seed;
} skw_shoot_params;
skw_status skw_photon_shoot(const skw_scene *scene, const skw_dl_scene *dl, const skw_shoot_params *params,
                            skw_photon_set **out, void *stream);
/* Photons (counts[3]), paths of each map (paths[3]) and paths shot (host outputs, any may be NULL). */
skw_status skw_photon_set_info(const skw_photon_set *set, uint64_t counts[3], uint64_t paths[3],
                               uint64_t *total_paths);
/* Host copies of map `which`: positions (3 floats), weights (3 floats), incident-direction and
 * normal ids (any may be NULL). */
skw_status skw_photon_set_export(const skw_photon_set *set, uint32_t which, float *pos, float *weight, uint16_t *dir,
                                 uint16_t *normal);
/* The photon map (KDTree<Photon>) of map `which`, built on the device (skw_photon_map_build). */
skw_status skw_photon_set_map(const skw_photon_set *set, uint32_t which, skw_photon_map **out, void *stream);
skw_status skw_photon_set_destroy(skw_photon_set *set);

/* ---------------------------------------------------------------- on-disk formats
 * Checkpoint / resume of the built structures, which the reference keeps only in memory: the scene
 * (flattened tree, triangles, instances, meshes, primitives), a photon map (kd-tree, payload) and
 * an irradiance photon map, each as a versioned little-endian file with an FNV-1a 64 trailer.
 * Loading skips every build and gives structures identical to the saved ones; files that fail the
 * magic, version, byte-order, size or checksum checks are rejected with SKW_INVALID_ARG. flags:
 * SKW_SCENE_HOST_ONLY loads without device copies. */
skw_status skw_scene_save(const skw_scene *scene, const char *path);
skw_status skw_scene_load(const char *path, uint32_t flags, skw_scene **out);
skw_status skw_photon_map_save(const skw_photon_map *map, const char *path);
skw_status skw_photon_map_load(const char *path, uint32_t flags, skw_photon_map **out);
skw_status skw_irradiance_map_save(const skw_irradiance_map *map, const char *path);
skw_status skw_irradiance_map_load(const char *path, skw_irradiance_map **out);

/* ---------------------------------------------------------------- PBRT scene import
 * Replaces PbrtSceneImporter::Import (Source/RayLibs/CADImport/SceneImporters/PbrtSceneImporter.h:
 * 31-40, PbrtSceneImporter_main.cpp, PbrtSceneImporter_hooks.cpp) for the part of a pbrt-v2 file
 * the traced path uses: trianglemesh shapes under the transform commands, attribute / transform
 * blocks, ReverseOrientation, matte and plastic materials (also named), diffuse area lights,
 * object instancing, Include, the perspective camera and the film resolution. Anything else is
 * skipped with a warning. Syntax errors (unterminated strings or lists, missing arguments, bad
 * numbers, unreadable files) fail with SKW_INVALID_ARG. */
typedef struct skw_pbrt_scene skw_pbrt_scene;

skw_status skw_pbrt_load(const char *path, skw_pbrt_scene **out);
/* Descriptions for skw_scene_create and skw_dl_scene_create (every primitive is its own material
 * group), the camera and the film size; any output may be NULL. The arrays stay owned by the
 * import and live until skw_pbrt_destroy. Fails when `cam` is asked for and no Camera was read. */
skw_status skw_pbrt_descs(const skw_pbrt_scene *scene, skw_scene_desc *desc, skw_dl_desc *dl, skw_camera *cam,
                          uint32_t *width, uint32_t *height);
skw_status skw_pbrt_warnings(const skw_pbrt_scene *scene, uint32_t *count);
skw_status skw_pbrt_warning(const skw_pbrt_scene *scene, uint32_t i, const char **message);
skw_status skw_pbrt_destroy(skw_pbrt_scene *scene);

/* ---------------------------------------------------------------- batched scenes
 * Many scenes in one persistent launch: every scene's camera rays (pixel-jittered as in
 * skw_dl_render, sample s of pixel p has index p*spp + s) are traced to their closest hit by one
 * kernel that takes ray chunks from a device work queue (smallest scenes first). The frame
 * (queue reset + trace) is captured once in a hipGraph and replayed. No counterpart in the
 * reference, which renders one scene at a time; its per-ray semantics are those of
 * skw_closest_hit. Scenes must not use instancing and need a traversal stack of at most 64. */
typedef struct skw_batch skw_batch;

typedef struct {
  skw_camera cam;
  uint32_t width, height, spp;
  uint64_t seed;
} skw_batch_view;

/* Launch shape: waves per workgroup (1, 2 or 4), LDS stack entries per lane (32 or 64, at least
 * the deepest scene's need), rays per queue chunk (a multiple of 64), resident waves per CU. */
typedef struct {
  uint32_t waves_per_block;
  uint32_t stack_entries;
  uint32_t chunk_rays;
  uint32_t waves_per_cu;
} skw_batch_config;

skw_status skw_batch_create(const skw_scene *const *scenes, const skw_batch_view *views, uint32_t n,
                            skw_batch **out);

/* A photon-gather job of a batch: skw_photon_irradiance's work (k nearest with PhotonFilter, disc
 * estimate) for nq query points, written to the job's own outputs every frame. */
typedef struct {
  const skw_photon_map *map;
  const double *d_points, *d_normals; /* 3 doubles per query, device */
  uint64_t nq;
  uint32_t k;
  double r;
  uint64_t paths;
  double rho;
  float *d_e_ext, *d_e_int;          /* 3 floats per query, device */
  float *d_radiance;                 /* optional */
} skw_batch_gather;

/* A batch whose queue also holds photon-gather jobs: ray chunks and query chunks of all scenes and
 * jobs are served by the same persistent launch (items ordered by size, smallest first). */
skw_status skw_batch_create_with_gathers(const skw_scene *const *scenes, const skw_batch_view *views, uint32_t n,
                                         const skw_batch_gather *gathers, uint32_t n_gathers, skw_batch **out);
skw_status skw_batch_destroy(skw_batch *batch);
/* Rebuilds the work queue for `cfg` and captures the frame graph. */
skw_status skw_batch_configure(skw_batch *batch, const skw_batch_config *cfg);
/* Times every valid launch shape (frames_per_config frames each, after one warm-up), keeps the
 * fastest (configured and captured on return) and reports it. */
skw_status skw_batch_autotune(skw_batch *batch, uint32_t frames_per_config, skw_batch_config *best,
                              double *best_ms);
/* Replays the captured frame on `stream`. */
skw_status skw_batch_frame(skw_batch *batch, void *stream);
/* Device pointer and count of a scene's hit records (valid after a frame). */
skw_status skw_batch_hits(const skw_batch *batch, uint32_t scene, const skw_hit **d_hits, uint64_t *n);
/* Completion time of each item's last chunk (scenes, then gather jobs), in nanoseconds after the
 * frame started (waits for the last frame). */
skw_status skw_batch_latency(skw_batch *batch, double *ns_per_scene);
/* Waits for the last frame and reports whether every wave of it ran to its end: SKW_HIP_ERROR if
 * a wave found an iteration in which none of its lanes could move and left the loop (a broken
 * invariant; the frame's hits are then incomplete). skw_batch_latency makes the same check.
 * No counterpart in the reference (its renderer has no persistent kernel). */
skw_status skw_batch_check(skw_batch *batch);
/* Writes scene `scene`'s camera rays (as the frame generates them) to d_rays; for checking. */
skw_status skw_batch_rays(const skw_batch *batch, uint32_t scene, skw_ray *d_rays, void *stream);
/* Total rays per frame. */
skw_status skw_batch_ray_count(const skw_batch *batch, uint64_t *n);

/* Stable sort of n (key, value) pairs of device uint32 arrays by key bits [begin_bit, end_bit):
 * the library's own radix sort, which orders the final gather's irradiance lookups and a shooting
 * round's photons. The inputs are not written; temporary memory comes from the library's pool.
 * No counterpart in the reference (its lookups run in pixel order on the host); exported so that
 * callers and tests can check it. */
skw_status skw_sort_pairs_u32(const uint32_t *d_keys_in, uint32_t *d_keys_out, const uint32_t *d_vals_in,
                              uint32_t *d_vals_out, uint64_t n, int begin_bit, int end_bit, void *stream);

/* Thread-local text of the last error. */
const char *skw_last_error(void);

/* Number of HIP devices visible (0 when there is none). */
int skw_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
