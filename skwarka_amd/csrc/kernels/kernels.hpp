// Launch interfaces of the gfx950 kernels (host side of the kernel translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../skw_layout.hpp"
#include "../../../include/skwarka_rt.h"

namespace skw {

struct TraceArgs {
  const DevNode *nodes;
  const DevTri *tris;
  const DevInst *insts;
  const uint32_t *inst_range;
  uint32_t root;
  const skw_ray *rays;
  skw_hit *hits;
  uint8_t *occluded;
  uint64_t n;
  const uint64_t *n_dev;  // when set, the ray count is read here on the device (n is ignored)
  uint64_t *counter;  // set to gridDim*blockDim before the trace kernel starts
  uint32_t *gstack;   // used only when the tree is deeper than the LDS stack
  uint32_t *pops;     // statistics variant only: node boxes tested per ray
  uint32_t *tests;    // statistics variant only: triangles tested per ray
  uint64_t *sums;     // statistics variant only: [0] += node boxes, [1] += triangles over all rays
  uint32_t refill_min;  // idle lanes a wave gathers before taking new rays (1..64)
  uint32_t leaf_min;    // lanes holding a leaf a wave gathers before testing leaves (1..64)
  uint32_t claim;       // rays a wave claims from the counter at once (64..4096)
  uint32_t nested;      // 1: the tree's child boxes lie inside their parents' (no NaN coordinate)
};

// Chooses the LDS stack size from stack_need (global-memory stack above 128 entries). wide: a
// walk without statistics on a stack of up to 32 entries runs the instance compiled for five waves
// per SIMD (the grid then holds more than 16 waves per CU).
hipError_t launch_closest_hit(const TraceArgs &args, bool inst, bool stats, uint32_t stack_need, uint32_t grid,
                              bool wide, hipStream_t stream);
hipError_t launch_any_hit(const TraceArgs &args, bool inst, bool stats, uint32_t stack_need, uint32_t grid, bool wide,
                          hipStream_t stream);
hipError_t launch_set_counter(uint64_t *counter, uint64_t value, hipStream_t stream);

struct GeomArgs {
  const DevPrim *prims;
  const DevMesh *meshes;
  const float *mesh_verts;
  const float *mesh_normals;
  const uint32_t *mesh_tris;
  const skw_ray *rays;
  skw_hit *hits;           // fix_bary: b1/b2 rewritten for primitives whose traversal values are not mesh-space
  skw_hit_geom *geom;      // hit_geometry output
  uint64_t n;
};
hipError_t launch_fix_bary(const GeomArgs &args, hipStream_t stream);
hipError_t launch_hit_geometry(const GeomArgs &args, hipStream_t stream);

// Per node: an outward-rounded box of the decoded normals of the photons in its subtree, as six
// int8 components (min x y z, max x y z) over 127, in one uint64 (subtrees of more than 2,048
// photons: the whole cube). Computed for the irradiance photon maps (built or loaded).
// With `near` (not null) it also writes each tree position's nearest-lookup record.
struct DevNearRec;
hipError_t launch_kd_nbox(const DevKdNode *nodes, const DevPhoton *photons, const double *cdir, uint64_t n,
                          uint64_t *nbox, DevNearRec *near, hipStream_t stream);

// A nearest lookup's view of one tree position of an irradiance map: the node (point, packed
// word), the photon's decoded normal (the direction table's doubles) and its subtree's normal box,
// in one 64-byte line, so that a visit makes one memory request and no dependent table read.
struct DevNearRec {
  float4 node;
  double nv[3];
  uint64_t nbox;
  uint64_t pad[2];
};
static_assert(sizeof(DevNearRec) == 64, "one 64-byte line per record");

struct GatherArgs {
  const DevKdNode *nodes;
  const DevPhoton *photons;
  const float *psplit;  // per node: its parent's split coordinate (0 for the root)
  const uint32_t *perm;
  const double *cdir;  // 65536 x 3
  uint64_t n_nodes;
  const double *points;
  const double *normals;
  uint64_t nq;
  uint32_t k;
  double r;
  double paths;
  double rho;
  double sin_t;         // sqrt(1 - 0.87^2), computed on the host
  int filter;           // 0 none, 1 PhotonFilter
  // heap storage: element (q, j) at q*q_stride + j*j_stride
  uint32_t *heap_pos;
  double *heap_d2;
  uint64_t q_stride, j_stride;
  // outputs (any may be null)
  uint32_t *ids;        // k per query (row-major): caller ids (ids_tree 0), tree positions (1) or the
  int ids_tree;         // photons' direction codes, the low 16 bits of their packed ids (2)
  double *d2;           // k per query
  uint32_t *count;
  float *e_ext, *e_int, *radiance;
  uint32_t *nearest;    // nearest mode
  uint32_t stack_need;
  uint32_t *gstack;
  // estimate over the k-nearest set: 0 disc (e_ext / e_int / radiance), 1 cone filter with
  // constant kc (e_ext), 2 Simpson-kernel caustic radiance toward wo (caustic, 3 doubles)
  int estimate;
  double kc;
  const double *wo;
  double *caustic;
  // optional visiting order: thread i serves query order[i] (its inputs and outputs); spatially
  // sorted orders keep a wave's lookups on nearby tree paths
  const uint32_t *order;
  // 1: a disc estimate runs its lookup first without writing the set's slots and repeats it with
  // them only when the set fills (set by the host when neither ids nor d^2 are returned)
  int defer;
  // optional per-node bounds of the subtree's photon normals (launch_kd_nbox); a nearest lookup
  // with the PhotonFilter does not walk a subtree none of whose photons can pass its normal test
  const uint64_t *nbox;
  // optional per-node records of the same map (launch_kd_nbox): nearest lookups read them instead
  // of the node, photon, normal-box and direction-table arrays
  const DevNearRec *near;
  // statistics instance: per query (visits, photons inside the radius, final set size)
  uint32_t *stats;
};
// mode 0: k nearest (+ estimate when e_ext != null), mode 1: nearest point. With args.stats the
// counting instance runs (same lookups, same outputs).
hipError_t launch_gather(const GatherArgs &args, int mode, hipStream_t stream);
// out[i] = (a[i] + b[i]) + c[i] in float over n values; a may be null (zeros).
// Photon-mapping composition per pixel: for every pass, ((direct + 0) + gather) with the renderer's
// sample rule, summed in pass order; rgb = float(sum / spp).
hipError_t launch_pm_combine(const double *sample_dl, const double *sample_fg, uint64_t npix, uint32_t spp,
                             float *rgb, hipStream_t stream);
hipError_t launch_irr_sum(const float *a, const float *b, const float *c, float *out, uint64_t n, hipStream_t stream);

// ------------------------------------------------------------------ direct lighting
struct DevMat {
  double kd[3];
  double ks[3];
  double exponent;  // Blinn exponent of a plastic material
  uint32_t type;    // 0 matte, 1 plastic
  uint32_t light;   // area light index of the primitive, 0xFFFFFFFF if none
  // parameter view (differentiable render): group, d kd / d reflectance (0 where clamped),
  // d exponent / d roughness (0 where clamped)
  uint32_t group;   // 0xFFFFFFFF if none
  uint32_t pad;
  double dkd;
  double de_dr;
};

struct DevLight {
  double radiance[3];
  double area;
  double intensity;
  double bmin[3], bmax[3];
  uint32_t tri_off, tri_count;
};

struct DevLightTri {
  double cdf;
  double n[3];  // GetTriangleNormal (float math) as double
  double v0[3], v1[3], v2[3];
};

struct DlArgs {
  GeomArgs geom;
  const DevMat *mats;        // sampling materials: BSDF sampling, pdfs, control flow
  const DevLight *lights;    // sampling lights: CDF intensities, radiance tests
  const DevLightTri *ltris;
  const DevMat *mats_e;      // evaluation materials: BSDF values
  const double *rad_e;       // 3 per light: evaluation radiance
  const double *rad_base;    // 3 per light: radiance at emission scale 1 (adjoint)
  uint32_t n_lights;
  uint32_t n_groups;
  double cam_o[3], cam_r[9], x_tan, y_tan;
  uint32_t width, height, spp, nls, nbs;
  uint64_t seed;
  uint32_t pass;      // sample index s of the launch's first pass
  uint32_t ppl_log2;  // passes per launch = 1 << ppl_log2: slot i is pixel slot i >> ppl_log2 of
                      // pass pass + (i & (passes - 1)), so a pixel's samples of the launch are adjacent
  uint64_t pix0;      // first pixel of this batch
  uint64_t batch;     // sample slots in this launch (pixels x passes per launch)
  skw_ray *prim_rays;
  skw_hit *prim_hits;
  double *le;         // 3 per sample: emitted radiance seen by the camera ray
  double *cdf;        // n_lights per sample
  double *ns;         // 3 per sample: shading normal
  double *point;      // 24 per sample: the shading point record of k_dl_point (dl.hip)
  skw_ray *shadow_rays;
  int32_t *shadow_idx;   // nls per sample: ray index or -1
  double *shadow_c;      // 3 per (sample, k): unoccluded contribution
  skw_ray *bsdf_rays;
  int32_t *bsdf_idx;     // nbs per sample
  double *bsdf_rec;      // 8 per (sample, k): f rgb, pdf, |wi.ns|/pdf, min_t, d f / d roughness, unused
  uint64_t *counts;      // [0] shadow rays, [1] BSDF rays of this pass (device)
  uint8_t *occluded;
  skw_hit *bsdf_hits;
  double *film;          // 3 per pixel, running sum over passes
  double *slot_L;        // several passes per launch: 4 per slot (L rgb, 1 when the sample is dropped),
                         // added to the film in pass order by k_dl_accum
  double *sample_L;      // when set: 3 per sample (pixel * spp + pass), Le + direct lighting before the
                         // renderer's sample rule, instead of the film (photon-mapping composition)
  uint64_t *totals;      // [0] primary, [1] shadow, [2] BSDF rays over the frame (device); [3..8]
                         // node / triangle tests of the three ray classes (counting frames)
  // adjoint (when adj is set): dLoss/dImage per pixel, gradient output, per-sample records
  const double *adj;     // 3 per pixel
  double *grad;          // 2 * n_groups + n_lights
  double *shadow_aux;    // 6 per (sample, k): f rgb (evaluation), weight, d f / d roughness, light
  int32_t *cam_light;    // per sample: light seen by the camera ray, or -1
  // tape (when tape_j is set): per sample of the frame (pass * pixels + pixel slot) a row of tape_stride
  // derivatives and the hit primitive's group
  double *tape_j;
  uint32_t *tape_group;
  uint32_t tape_stride;
};

struct DlTapeArgs {
  const double *tape_j;
  const uint32_t *tape_group;
  uint32_t stride, n_groups, n_lights, spp, width, height;
  uint64_t batch, n_samples;
  const double *adj;
  double *grad;
};
hipError_t launch_dl_tape_adjoint(const DlTapeArgs &t, hipStream_t stream);
constexpr uint32_t kDlMaxGroups = 64;
constexpr uint32_t kDlMaxParams = 2 * kDlMaxGroups + 16;

// Photon final gather over the camera hits of a direct-lighting pass (dl.hip).
constexpr uint32_t kFgBlock = 64;  // slots per block of the gather rays' layout
struct FgArgs {
  DlArgs a;                       // camera, scene, materials, shading-point records, film
  uint32_t gs, gs_sqrt;           // gather samples (a square) and its side
  const double *cdir;             // 65536 x 3 direction table
  const uint32_t *near_ids;       // 32 per slot: the nearest indirect photons' direction codes
  const uint32_t *near_count;     // per slot (null: no indirect photons)
  double *s_pts, *s_nrm;          // 3 each per slot: shading point / geometric normal for the lookup
  skw_ray *rays;                  // gather rays: 2 gs per slot (BSDF samples, then cone samples), sample-major
                                  // within blocks of kFgBlock slots (dl.hip fg_ray)
  uint64_t slots_padded;          // slots of this batch rounded up to kFgBlock
  uint32_t *plastic;              // gather rays whose hit is plastic (k_fg_hit lists them for k_fg_plastic)
  uint64_t *plastic_count;        // their number (device)
  double hole_o[3], hole_d[3];    // a ray outside the scene box pointing away (samples that cast none)
  double *w;                      // 3 per gather ray: its weight
  uint64_t *count;                // gather rays cast in this batch (device)
  uint64_t max_rays;              // slots * 2 gs
  skw_hit *hits;                  // gather hits
  double *q_pts, *q_nrm;          // 3 each per gather ray: irradiance lookup point / normal (NaN when missed)
  double *ts;                     // 3 per gather ray: total reflection scattering at the hit
  uint8_t *side;                  // per gather ray: 1 when the gather direction is on the normal's side
  const uint32_t *irr_id;         // per gather ray: nearest irradiance photon or 0xFFFFFFFF
  const float *e_ext, *e_int;     // irradiance photons' irradiance (creation order)
  double *radiance;               // 3 per sample (pixel * spp + s), optional
  uint32_t *keys, *order;         // per gather ray: cell key of its lookup point (misses last) and its index
  double lo[3], cell[3];          // scene box corner and cells per unit length (2^morton_bits per axis)
  uint32_t morton_bits;           // bits per axis of the lookup points' Morton cells
  // The plastic lobe's 5 x 5 quadrature (k_fg_quad): quad[0..4] cos and quad[5..9] sin of the
  // five azimuths, then per material m quad[10 + 10 m ..] the five cos_h and quad[15 + 10 m ..]
  // the five sin_h of its Blinn exponent.
  double *quad;
  uint32_t n_mats;
};
hipError_t launch_fg_lookup_in(const FgArgs &f, hipStream_t stream);
hipError_t launch_fg_samples(const FgArgs &f, hipStream_t stream);
hipError_t launch_fg_hit(const FgArgs &f, hipStream_t stream);
hipError_t launch_fg_quad(const FgArgs &f, hipStream_t stream);
hipError_t launch_fg_resolve(const FgArgs &f, hipStream_t stream);
// Photon shooting, one round of paths (shoot.hip). Active paths live in ping-pong lists: bounce
// kernels read list `cur` (count counts[cur]) and append survivors to the other list.
struct ShootPath {
  double w[3];      // photon weight
  uint32_t path;    // path index
  uint32_t n_isect; // intersections so far
};
struct ShootArgs {
  GeomArgs geom;
  const DevMat *mats;
  const DevLight *lights;
  const DevLightTri *ltris;
  const double *light_cdf;  // power CDF over the lights
  uint32_t n_lights;
  uint64_t first, n;        // the round's paths [first, first + n)
  uint64_t seed;
  uint32_t done[3];         // maps (caustic, direct, indirect) that take no photons this round
  uint32_t cur;             // list read by the bounce kernel (0 or 1)
  skw_ray *rays[2];
  ShootPath *state[2];
  const skw_hit *hits;      // closest hits of list `cur`
  uint64_t *counts;         // [0], [1] list sizes; [2 + m] photons of map m this round
  float *ph_pos[3], *ph_w[3];
  uint32_t *ph_codes[3];    // incident direction id | normal id << 16
  uint32_t *ph_key[3];      // path - first (the photons are later ordered by it, stably)
};
hipError_t launch_shoot_start(const ShootArgs &a, hipStream_t stream);
// Grid for n_bound paths; the kernel reads the exact count of list `cur` from counts[cur].
hipError_t launch_shoot_bounce(const ShootArgs &a, uint64_t n_bound, hipStream_t stream);
hipError_t launch_iota_u32(uint32_t *out, uint64_t n, hipStream_t stream);  // out[j] = j
// out[j] = in[order[j]] for the photon records of one map (n records).
hipError_t launch_photon_gather_records(const float *pos_in, const float *w_in, const uint32_t *codes_in,
                                        const uint32_t *order, uint64_t n, float *pos_out, float *w_out,
                                        uint32_t *codes_out, hipStream_t stream);

// Photon kd-tree built on the device (kd_build.hip), element for element the tree of
// KDTree::_Build with libstdc++'s nth_element.
struct KdBuildArgs {
  void *items;           // n kd items (position + caller index, 16 bytes each)
  uint32_t *word;        // n node words
  uint32_t *scr_l, *scr_r;  // n each: partition scratch of the large nodes
  uint64_t n;
};
constexpr uint64_t kKdLdsItems = 2048;  // nodes up to this size are built in LDS
constexpr uint64_t kKdSeqMax = 32;      // subtrees up to this size are built by one thread
hipError_t launch_kd_init(const KdBuildArgs &a, const float *pos, hipStream_t stream);
// One level's nodes: ranges as (begin, end) pairs; big = longer than kKdLdsItems.
hipError_t launch_kd_big(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, hipStream_t stream);
hipError_t launch_kd_mid(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, uint64_t max_size,
                         hipStream_t stream);
hipError_t launch_kd_small(const KdBuildArgs &a, const uint64_t *ranges, uint32_t n_ranges, hipStream_t stream);
// Large nodes of the top levels, many workgroups per node: the state of one node's introselect.
struct KdStepNode {
  int64_t b, e, f, l, nth, depth;
  int64_t f_old, K;        // the last step's start and swap count
  int32_t axis, active;
  float pivot;
  uint32_t nl, nr;         // the last step's candidate counts (>= pivot from the left, <= pivot)
  uint32_t chunk0, nchunks;  // this node's chunks in the level's chunk list
};
constexpr uint64_t kKdChunk = 16384;  // positions per workgroup of the multi-workgroup steps
// chunk_node[j]: node of chunk j; work: 8 floats + 4 uint32 per chunk.
hipError_t launch_kd_step_init(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, const uint32_t *chunk_node,
                               uint32_t n_chunks, float *work, hipStream_t stream);
// One introselect step: begin (median of three, pivot; nodes leave when short or out of depth).
hipError_t launch_kd_step_begin(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, hipStream_t stream);
hipError_t launch_kd_step_rest(const KdBuildArgs &a, KdStepNode *nodes, uint32_t n_nodes, const uint32_t *chunk_node,
                               uint32_t n_chunks, float *work, hipStream_t stream);
// The rest of each node in one workgroup (LDS tail, sequential end, node word).
hipError_t launch_kd_step_tail(const KdBuildArgs &a, const KdStepNode *nodes, uint32_t n_nodes, hipStream_t stream);
// Tree-order node records, permutation, payload (weights; codes = dir | normal << 16 by caller
// index, or the two 16-bit id arrays) and parent splits.
hipError_t launch_kd_finish(const KdBuildArgs &a, const float *weight, const uint32_t *codes, const uint16_t *dir,
                            const uint16_t *normal, DevKdNode *nodes, DevPhoton *photons, uint32_t *perm,
                            float *psplit, hipStream_t stream);

// Sorts n (key, value) pairs by key bits [begin_bit, end_bit) (sort.hip). With temp == nullptr
// only sets temp_bytes.
hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                          const uint32_t *vals_in, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit,
                          hipStream_t stream);

hipError_t launch_dl_clear(const DlArgs &a, hipStream_t stream);
hipError_t launch_dl_camera(const DlArgs &a, hipStream_t stream);
hipError_t launch_dl_point(const DlArgs &a, hipStream_t stream);  // shading-point records only
hipError_t launch_dl_shade(const DlArgs &a, hipStream_t stream);
hipError_t launch_dl_resolve(const DlArgs &a, hipStream_t stream);
hipError_t launch_dl_finish(const DlArgs &a, float *rgb, hipStream_t stream);
hipError_t launch_dl_accum(const DlArgs &a, hipStream_t stream);
hipError_t launch_dl_finish_f64(const DlArgs &a, double *img, hipStream_t stream);
hipError_t launch_dl_clear_grad(double *grad, uint32_t n, hipStream_t stream);

// Materials and lights of one parameter vector, built on the device from the scene's base tables
// (the vectors travel as kernel arguments, so no host buffer has to outlive the call).
struct DlParamArgs {
  const DevMat *base_mats;
  const DevLight *base_lights;
  uint32_t n_prims, n_lights, n_groups;
  DevMat *mats_s, *mats_e;
  DevLight *lights_s;
  double *rad_e, *rad_base;
  double theta_s[kDlMaxParams];
  double theta_e[kDlMaxParams];
};
hipError_t launch_dl_params(const DlParamArgs &a, hipStream_t stream);

// ------------------------------------------------------------------ batched scenes
struct DevBatchScene {
  const DevNode *nodes;
  const DevTri *tris;
  uint32_t root;
  uint32_t width, height, spp;
  double cam_o[3], cam_r[9], x_tan, y_tan;
  uint64_t seed;
  uint64_t n_rays;      // width * height * spp
  uint64_t hit_offset;  // first record of this scene in the hit buffer
  uint32_t nested;      // 1: the tree's child boxes lie inside their parents' (no NaN coordinate)
  uint32_t pad;
};

// A queue chunk: a ray range of a scene, or (item >= n_scenes) a query range of gather job
// item - n_scenes.
struct DevChunk {
  uint32_t item;
  uint32_t pad;
  uint64_t begin, end;
};

struct BatchArgs {
  const DevBatchScene *scenes;
  const GatherArgs *jobs;  // photon gathers (disc estimate), one record per job
  const DevChunk *chunks;
  uint64_t n_chunks;
  uint64_t *queue_head;
  uint64_t *item_done;   // rays / queries finished per item (scenes, then jobs)
  uint64_t *item_end;    // wall clock of each item's completion
  uint64_t *frame_start; // wall clock at frame start
  uint64_t *stalled;     // set to 1 by a wave that found an iteration in which no lane could move
  uint32_t n_scenes;
  uint32_t n_jobs;
  skw_hit *hits;
};
hipError_t launch_batch_reset(const BatchArgs &a, hipStream_t stream);
// waves: 1, 2 or 4 per workgroup; stack: 32 or 64 entries per lane.
hipError_t launch_batch_trace(const BatchArgs &a, uint32_t waves, uint32_t stack, uint32_t grid, hipStream_t stream);
hipError_t launch_batch_rays(const DevBatchScene *scene, skw_ray *rays, uint64_t n, hipStream_t stream);

// Device build of the triangle tree (one level of nodes per launch pair). Node records are
// allocated in blocks of siblings; info = flags (0 leaf, axis + 1) | child mask << 2 |
// split-axis mask << 5 | depth << 16.
struct TriBuildNode {
  float bmin[3], bmax[3];
  uint32_t tb, te;        // triangle range (positions in the partition arrays)
  uint32_t first_child;   // id of the first present child, 0xFFFFFFFF for a leaf
  uint32_t info;
};
struct TriBuildArgs {
  float *box;     // 6 per position: triangle box min xyz, max xyz
  uint32_t *ref;  // per position: triangle id
  float *box2;    // scatter targets of the partition (same sizes)
  uint32_t *ref2;
  // per-position partition scratch (nine arrays of the triangle count)
  uint32_t *s_cls, *s_rf, *s_nb, *s_G, *s_Rpos, *s_order, *s_elem, *s_ptr, *s_dest;
  TriBuildNode *nodes;
  uint32_t *node_count;  // ids handed out
  const uint32_t *list;  // nodes of this launch
  uint32_t list_n;
  uint32_t *next_big, *next_big_n;      // children larger than small_max
  uint32_t *next_small, *next_small_n;  // children of at most small_max triangles
  uint32_t *next_wide, *next_wide_n;    // children larger than wide_min
  uint32_t small_max;                   // at most 64
  uint32_t wide_min;
  uint32_t node_cap;                    // node records allocated (subtree builds check it)
  uint32_t *overflow;                   // set when a subtree build ran out of node records
};
// k_tb_level: one 256-thread workgroup per listed node, children listed by size;
// launch_tri_build_small: one wave per listed node builds its whole subtree.
hipError_t launch_tri_build_level(const TriBuildArgs &a, hipStream_t stream);
hipError_t launch_tri_build_small(const TriBuildArgs &a, hipStream_t stream);

// World-space triangles of a scene without instancing, in the tree builder's order (meshes used by
// one primitive, in mesh order; their cleaned triangles in order), moved by the primitive's
// mesh-to-world matrix as the host builder does (Transform::operator() on float points), with
// their float boxes and the identity order.
struct TriPrepArgs {
  const DevMesh *meshes;
  const float *verts;
  const uint32_t *tris;
  const DevPrim *prims;
  const uint64_t *goff;  // n_meshes + 1: first triangle of each mesh in the order (0 triangles if unused)
  const uint32_t *user;  // per mesh: its primitive
  uint32_t n_meshes;
  uint64_t n;
  float *tv;             // 9 per triangle
  float *box;            // 6 per triangle
  uint32_t *tprim, *ttri, *ref;
  uint32_t *flags;       // bit 0: a box with a non-finite coordinate
};
hipError_t launch_tri_prep(const TriPrepArgs &a, hipStream_t stream);

// The host builder's depth-first layout of a device-built tree (children blocks appended as nodes
// are placed), computed level by level from subtree sizes, and the triangle array in final order.
struct TriLayoutArgs {
  const TriBuildNode *hn;
  uint32_t count;        // node records
  uint32_t n;            // triangles
  uint32_t *hist;        // per depth (kTriLayoutDepths): nodes
  uint32_t *fill;        // per depth: scatter cursor
  const uint32_t *level_off;  // per depth + 1: first entry of the depth in by_depth
  uint32_t *by_depth;    // node ids grouped by depth
  uint32_t *size;        // subtree node counts
  uint32_t *cval;        // children of the nodes before it in depth-first order
  uint32_t *idx;         // layout index
  uint32_t *flags;       // bit 1: a record out of range
  uint32_t *leaves;
  const float *tv;
  const uint32_t *tprim, *ttri, *ref;
  DevNode *nodes;
  DevTri *tris;
  uint32_t level_begin, level_end;  // slice of by_depth of one level launch
};
constexpr uint32_t kTriLayoutDepths = 256;
hipError_t launch_tri_layout_hist(const TriLayoutArgs &a, hipStream_t stream);     // hist, leaves, flags
hipError_t launch_tri_layout_scatter(const TriLayoutArgs &a, hipStream_t stream);  // by_depth, root's idx / cval
hipError_t launch_tri_layout_size(const TriLayoutArgs &a, hipStream_t stream);     // one level, bottom-up
hipError_t launch_tri_layout_place(const TriLayoutArgs &a, hipStream_t stream);    // one level, top-down
hipError_t launch_tri_layout_tris(const TriLayoutArgs &a, hipStream_t stream);

// The largest nodes of a level, each split over chunks of `chunk` triangles (one workgroup each).
struct TriWide {
  uint32_t id, tb, n, info;   // node id, range start, triangles, info bits of the node record
  uint32_t c_begin, c_end;    // its chunks
  uint32_t leaf, axis;
  double split;
  unsigned long long key[6];  // node box keys (first extremes)
  uint32_t n_right, n_nonr, K, n_left, m, r_total, pad[2];
};
struct TriWideArgs {
  TriBuildArgs b;          // list / list_n: the wide nodes
  TriWide *w;
  void *bins;              // 3 bin sets per wide node (tri_build_bins_bytes() each)
  uint32_t *chunk_w;       // per chunk: wide node index
  uint32_t *ccnt, *cbase;  // per chunk: 8 counters, 8 prefix bases
  uint32_t *info_out;      // [0] chunks, [1] largest node
  uint32_t chunk_cap;      // chunk table entries allocated
  uint32_t n_chunks;
  uint32_t chunk;          // triangles per chunk (multiple of 256)
};
size_t tri_build_bins_bytes();
// setup: the chunk table and per-node state from the listed nodes (info_out read by the host);
// then the build of the listed nodes' children with jump_rounds pointer-jumping rounds.
hipError_t launch_tri_wide_setup(const TriWideArgs &a, hipStream_t stream);
hipError_t launch_tri_wide_level(const TriWideArgs &a, uint32_t jump_rounds, hipStream_t stream);

}  // namespace skw
