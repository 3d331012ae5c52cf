// Direct-lighting wavefront on gfx950: camera rays -> closest hit -> shade (emission, light and
// BSDF sampling, ray emission) -> occlusion / closest-hit of the emitted rays -> resolve (MIS
// weights of BSDF rays that reach a light, per-sample sum, film accumulation).
//
// Per-sample arithmetic follows:
//   DirectLightingIntegrator::ComputeDirectLighting / _SampleLights / _SampleBSDF
//       (Source/RayLibs/Raytracer/Core/DirectLightingIntegrator.cpp:54-366)
//   BSDF::Evaluate / PDF / Sample and its shading frame (Raytracer/Core/BSDF.cpp:22-200, BSDF.h:182-244)
//   BxDF::Sample / PDF, Lambertian (Raytracer/Core/BxDF.cpp:20-41, BxDFs/Lambertian.cpp:18-31)
//   Microfacet<FresnelDielectric, BlinnDistribution> (BxDFs/Microfacet.h:89-152,
//       MicrofacetDistributions/BlinnDistribution.h:70-110, Core/Fresnel.h:92-124)
//   DiffuseAreaLightSource::SampleLighting / LightingPDF / Radiance (LightSources/DiffuseAreaLightSource.cpp:57-174)
//   IrradianceLightsSamplingStrategy::GetLightsCDF (LightsSamplingStrategies/IrradianceLightsSamplingStrategy.cpp:54-129)
//   MathRoutines::SubtendedSolidAngle / BinarySearchCDF; SamplingRoutines::ConcentricDiskSampling /
//       PowerHeuristic (Source/RayLibs/Math/MathRoutines.h:192-215, 319-390; SamplingRoutines.h:197-325, 449-455)
//   PerspectiveCamera::GenerateRay (Raytracer/Cameras/PerspectiveCamera.cpp:36-64)
//   the NaN / negative / infinite sample rule (Raytracer/Renderers/SamplerBasedRenderer.cpp:458-478).
// Rays go through the scene's own traversal kernels. Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "shade_device.hpp"

namespace skw {
namespace {

using namespace dev;

// Pixel of sample slot i of a pass: 8x8 pixel tiles in row order, so the 64 camera rays of a wave
// (and the occlusion / BSDF rays compacted from them in slot order) start from one tile. The
// sample stream depends only on the pixel, so the mapping changes no value. Images whose sides are
// not multiples of 8 use row order.
__device__ __forceinline__ uint64_t pixel_of(uint32_t width, uint32_t height, uint64_t s) {
  if ((width & 7u) != 0 || (height & 7u) != 0) return s;
  const uint64_t tile = s >> 6, in = s & 63u, tiles_x = width >> 3;
  const uint64_t x = (tile % tiles_x) * 8 + (in & 7u), y = (tile / tiles_x) * 8 + (in >> 3);
  return y * width + x;
}
// Sample slot i of a launch: pixel slot i >> ppl_log2, pass pass + (i & (passes - 1)). A pixel's
// samples of the launch are adjacent slots, so a wave's 64 camera rays cover 64 >> ppl_log2 pixels
// of one tile (with one pass per launch, 64 pixels of one pass).
__device__ __forceinline__ uint64_t pixel_of(const DlArgs &a, uint64_t i) {
  return pixel_of(a.width, a.height, a.pix0 + (i >> a.ppl_log2));
}
__device__ __forceinline__ uint32_t pass_of(const DlArgs &a, uint64_t i) {
  return a.pass + (uint32_t)(i & ((1ull << a.ppl_log2) - 1ull));
}
// The sample's row of the frame's tape: pass-major over the batch's pixel slots.
__device__ __forceinline__ uint64_t tape_row(const DlArgs &a, uint64_t i) {
  return (uint64_t)pass_of(a, i) * (a.batch >> a.ppl_log2) + (i >> a.ppl_log2);
}

// ------------------------------------------------------------------ kernels

__global__ __launch_bounds__(256) void k_dl_camera(DlArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    a.counts[0] = 0;
    a.counts[1] = 0;
  }
  if (i >= a.batch) return;
  const uint64_t pix = pixel_of(a, i);
  const uint32_t px = (uint32_t)(pix % a.width), py = (uint32_t)(pix / a.width);
  const uint64_t sidx = pix * a.spp + pass_of(a, i);
  double ix = (double)px + rnd(a.seed, sidx, 0), iy = (double)py + rnd(a.seed, sidx, 1);
  double cx = 2.0 * a.x_tan * (ix - a.width * 0.5) / a.width;
  double cy = 2.0 * a.y_tan * (iy - a.height * 0.5) / a.height;
  D3 d{a.cam_r[0] * cx + a.cam_r[1] * cy + a.cam_r[2], a.cam_r[3] * cx + a.cam_r[4] * cy + a.cam_r[5],
       a.cam_r[6] * cx + a.cam_r[7] * cy + a.cam_r[8]};
  normalize3(d);
  put_ray(a.prim_rays + i, D3{a.cam_o[0], a.cam_o[1], a.cam_o[2]}, d, 0.0, __builtin_huge_val());
}

// Shading is split in three kernels so that each keeps few registers (more resident waves hide
// the dependent loads): k_dl_point (one thread per camera sample: hit geometry, emitted radiance,
// the light CDF, the BSDF frame), then k_dl_light and k_dl_bsdf (one thread per (sample, k):
// light sample k / BSDF sample k of _SampleLights / _SampleBSDF, and the ray it casts). The
// shading point travels through a per-sample record of kPointDoubles doubles:
//   p[0..2] wo[3..5] ns[6..8] ng[9..11] e1[12..14] e2[15..17] cross[18..20] dot[21] prim bits[22].
constexpr uint32_t kPointDoubles = 24;

__global__ __launch_bounds__(64) void k_dl_point(DlArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.batch) return;
  const skw_hit h = a.prim_hits[i];
  const bool hit = h.prim != kMiss;
  const uint32_t nl = a.n_lights;
  double le[3] = {0.0, 0.0, 0.0};
  int32_t cam_light = -1;
  double rec[kPointDoubles];
  for (uint32_t j = 0; j < kPointDoubles; ++j) rec[j] = 0.0;
  rec[22] = __longlong_as_double((long long)kMiss);
  if (hit) {
    const uint32_t prim = h.prim;
    const skw_ray &r = a.prim_rays[i];
    const D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
    const HitGeom g = hit_geometry(a.geom, o, d, prim, h.tri);
    const D3 wo = mul3(d, -1.0);
    const DevMat &m = a.mats[prim];
    if (m.light != kMiss && dot3(wo, g.ng) > 0.0) {
      for (int c = 0; c < 3; ++c) le[c] = a.rad_e[3 * m.light + c];
      cam_light = (int32_t)m.light;
    }
    // BSDF frame: DG tangent of a mesh without UVs or tangents (TriangleMesh.cpp:520-540)
    const D3 ns = g.ns;
    D3 tangent, e2;
    bsdf_frame(ns, tangent, e2);
    if (nl > 0) {
      D3 nrm = dot3(wo, ns) >= 0.0 ? ns : mul3(ns, -1.0);
      double cdf[kMaxLights];
      lights_cdf(a, g.p, nrm, cdf);
      for (uint32_t l = 0; l < nl; ++l) a.cdf[i * nl + l] = cdf[l];
    }
    const D3 vs[7] = {g.p, wo, ns, g.ng, tangent, e2, g.cross};
    for (int v = 0; v < 7; ++v) {
      rec[3 * v] = vs[v].x;
      rec[3 * v + 1] = vs[v].y;
      rec[3 * v + 2] = vs[v].z;
    }
    rec[21] = g.dot;
    rec[22] = __longlong_as_double((long long)prim);
    a.ns[3 * i] = ns.x;
    a.ns[3 * i + 1] = ns.y;
    a.ns[3 * i + 2] = ns.z;
  }
  double2 *out = reinterpret_cast<double2 *>(a.point + i * kPointDoubles);
  for (uint32_t j = 0; j < kPointDoubles / 2; ++j) out[j] = make_double2(rec[2 * j], rec[2 * j + 1]);
  for (int c = 0; c < 3; ++c) a.le[3 * i + c] = le[c];
  if (a.cam_light) a.cam_light[i] = cam_light;
}

// The shading point of sample i: position, outgoing direction, the BSDF with the sampling
// material (b) and the evaluation material (be, when SPLIT or ADJ), the primitive, m_cross / m_dot.
struct Point {
  D3 p, wo;
  HitGeom g;
  uint32_t prim;
};

__device__ __forceinline__ bool load_point(const DlArgs &a, uint64_t i, Point &pt) {
  const double2 *q = reinterpret_cast<const double2 *>(a.point + i * kPointDoubles);
  double v[kPointDoubles];
  for (uint32_t j = 0; j < kPointDoubles / 2; ++j) {
    const double2 x = q[j];
    v[2 * j] = x.x;
    v[2 * j + 1] = x.y;
  }
  pt.prim = (uint32_t)__double_as_longlong(v[22]);
  pt.p = D3{v[0], v[1], v[2]};
  pt.wo = D3{v[3], v[4], v[5]};
  pt.g.p = pt.p;
  pt.g.ns = D3{v[6], v[7], v[8]};
  pt.g.ng = D3{v[9], v[10], v[11]};
  pt.g.cross = D3{v[18], v[19], v[20]};
  pt.g.dot = v[21];
  return pt.prim != kMiss;
}

__device__ __forceinline__ void point_bsdf(const DlArgs &a, uint64_t i, uint32_t prim, Bsdf &b) {
  const double2 *q = reinterpret_cast<const double2 *>(a.point + i * kPointDoubles);
  const double2 n01 = q[3], n2g0 = q[4], g12 = q[5], t01 = q[6], t2u0 = q[7], u12 = q[8];
  b.n = D3{n01.x, n01.y, n2g0.x};
  b.ng = D3{n2g0.y, g12.x, g12.y};
  b.e1 = D3{t01.x, t01.y, t2u0.x};
  b.e2 = D3{t2u0.y, u12.x, u12.y};
  const DevMat &m = a.mats[prim];
  b.plastic = m.type == 1;
  for (int c = 0; c < 3; ++c) {
    b.kd[c] = m.kd[c];
    b.ks[c] = m.ks[c];
  }
  b.e = m.exponent;
}

// _SampleLights, light sample k of camera sample i: one occlusion ray when the unoccluded
// contribution is not black. SPLIT: evaluation materials differ from sampling materials
// (replayed estimator). ADJ: also record what the adjoint needs (d f / d roughness, weight, light).
template <bool SPLIT, bool ADJ>
__global__ __launch_bounds__(kSampleBlock) void k_dl_light(DlArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nls = a.nls, nl = a.n_lights;
  const uint64_t i = t / nls;
  const uint32_t k = (uint32_t)(t - i * nls);
  const bool in = i < a.batch;
  Point pt;
  const bool shade = in && nl > 0 && load_point(a, i, pt);
  bool want = false;
  D3 wi{0, 0, 0};
  double lf[3] = {0, 0, 0};
  double len = 0.0, mint = 0.0;
  double fe[3] = {0, 0, 0}, w = 0.0, dfr = 0.0;
  uint32_t l = 0;
  if (shade) {
    Bsdf b;
    point_bsdf(a, i, pt.prim, b);
    const uint64_t sidx = pixel_of(a, i) * a.spp + pass_of(a, i);
    const double *cdf = a.cdf + i * nl;
    double uc = rnd(a.seed, sidx, 2 + k);
    double u0 = rnd(a.seed, sidx, 6 + 2 * k), u1 = rnd(a.seed, sidx, 7 + 2 * k);
    l = search_cdf([&](uint32_t j) { return cdf[j]; }, nl, uc);
    double comp_pdf = l == 0 ? cdf[0] : cdf[l] - cdf[l - 1];
    const DevLight &L = a.lights[l];
    double tri_u = (l == 0 ? uc : uc - cdf[l - 1]) / comp_pdf;
    if (L.tri_count > 0) {
      const DevLightTri *lt = a.ltris + L.tri_off;
      uint32_t ti = search_cdf([&](uint32_t j) { return lt[j].cdf; }, L.tri_count, tri_u);
      const DevLightTri &T = lt[ti];
      double tmp = sqrt(u0);
      double bb1 = 1.0 - tmp, bb2 = u1 * tmp;
      D3 v0{T.v0[0], T.v0[1], T.v0[2]}, v1{T.v1[0], T.v1[1], T.v1[2]}, v2{T.v2[0], T.v2[1], T.v2[2]};
      D3 sp = add3(add3(mul3(v0, bb1), mul3(v1, bb2)), mul3(v2, 1.0 - bb1 - bb2));
      D3 lv = sub3(sp, pt.p);
      len = sqrt(len_sqr3(lv));
      wi = mul3(lv, 1.0 / len);
      D3 ln{T.n[0], T.n[1], T.n[2]};
      double ac = fabs(dot3(wi, ln));
      double lpdf = ac > kEps ? (len * len) / (L.area * ac) : 0.0;
      bool lit = dot3(ln, wi) < 0.0;
      double lr[3] = {lit ? L.radiance[0] : 0.0, lit ? L.radiance[1] : 0.0, lit ? L.radiance[2] : 0.0};
      if (lpdf > 0.0 && !(lr[0] == 0.0 && lr[1] == 0.0 && lr[2] == 0.0)) {
        double f[3];
        // ADJ runs at one parameter vector (sampling and evaluation materials agree); d exponent /
        // d roughness comes from the evaluation table.
        if (ADJ) b.evaluate<true>(pt.wo, wi, f, a.mats_e[pt.prim].de_dr, dfr);
        else b.evaluate(pt.wo, wi, f);
        for (int c = 0; c < 3; ++c) lf[c] = lr[c] * f[c];
        if (!(lf[0] == 0.0 && lf[1] == 0.0 && lf[2] == 0.0)) {
          mint = next_min_t(a.geom.prims[pt.prim], pt.g, wi);
          double bp = b.pdf(pt.wo, wi);
          w = power_heuristic((double)a.nls, lpdf * comp_pdf, (double)a.nbs, bp);
          w *= fabs(dot3(wi, b.n)) / (lpdf * comp_pdf);
          for (int c = 0; c < 3; ++c) fe[c] = f[c];
          if (SPLIT) {
            Bsdf be = b;
            const DevMat &me = a.mats_e[pt.prim];
            for (int c = 0; c < 3; ++c) {
              be.kd[c] = me.kd[c];
              be.ks[c] = me.ks[c];
            }
            be.e = me.exponent;
            be.evaluate(pt.wo, wi, fe);
            for (int c = 0; c < 3; ++c) lf[c] = (lit ? a.rad_e[3 * l + c] : 0.0) * fe[c];
          } else {
            for (int c = 0; c < 3; ++c) lf[c] = (lit ? a.rad_e[3 * l + c] : 0.0) * f[c];
          }
          for (int c = 0; c < 3; ++c) lf[c] = lf[c] * w;
          want = true;
        }
      }
    }
  }
  const int32_t slot = append_slot_block(a.counts, want);
  if (in) a.shadow_idx[i * nls + k] = slot;
  if (want) {
    put_ray(a.shadow_rays + slot, pt.p, wi, mint, len - 1e-4);
    double *c3 = a.shadow_c + 3 * (i * nls + k);
    c3[0] = lf[0];
    c3[1] = lf[1];
    c3[2] = lf[2];
    if (ADJ) {
      double *x = a.shadow_aux + 6 * (i * nls + k);
      x[0] = fe[0];
      x[1] = fe[1];
      x[2] = fe[2];
      x[3] = w;
      x[4] = dfr;
      x[5] = (double)l;
    }
  }
}

// _SampleBSDF, BSDF sample k of camera sample i: one closest-hit ray when the pdf is positive and
// the value not black.
template <bool SPLIT, bool ADJ>
__global__ __launch_bounds__(kSampleBlock) void k_dl_bsdf(DlArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nbs = a.nbs;
  const uint64_t i = t / nbs;
  const uint32_t k = (uint32_t)(t - i * nbs);
  const bool in = i < a.batch;
  Point pt;
  const bool shade = in && a.n_lights > 0 && load_point(a, i, pt);
  bool want = false;
  D3 wi{0, 0, 0};
  double f[3] = {0, 0, 0}, bp = 0.0, mint = 0.0, dfr = 0.0, cosw = 0.0;
  if (shade) {
    Bsdf b;
    point_bsdf(a, i, pt.prim, b);
    const uint64_t sidx = pixel_of(a, i) * a.spp + pass_of(a, i);
    double uc = rnd(a.seed, sidx, 14 + k);
    double u0 = rnd(a.seed, sidx, 18 + 2 * k), u1 = rnd(a.seed, sidx, 19 + 2 * k);
    D3 wo_l, wi_l;
    b.sample(pt.wo, u0, u1, uc, wi, bp, f, wo_l, wi_l);
    if (bp > 0.0 && !(f[0] == 0.0 && f[1] == 0.0 && f[2] == 0.0)) {
      mint = next_min_t(a.geom.prims[pt.prim], pt.g, wi);
      cosw = fabs(dot3(wi, b.n)) / bp;
      want = true;
      // f is non-black, so the sampled pair passed the geometric-normal test: re-evaluate
      // the same local pair with the evaluation materials.
      if (SPLIT || ADJ) {
        Bsdf be = b;
        const DevMat &me = a.mats_e[pt.prim];
        for (int c = 0; c < 3; ++c) {
          be.kd[c] = me.kd[c];
          be.ks[c] = me.ks[c];
        }
        be.e = me.exponent;
        if (SPLIT) be.eval_reflect_local(wo_l, wi_l, f);
        if (ADJ) be.eval_reflect_local<true>(wo_l, wi_l, f, me.de_dr, dfr);
      }
    }
  }
  const int32_t slot = append_slot_block(a.counts + 1, want);
  if (in) a.bsdf_idx[i * nbs + k] = slot;
  if (want) {
    put_ray(a.bsdf_rays + slot, pt.p, wi, mint, __builtin_huge_val());
    double *rec = a.bsdf_rec + 8 * (i * nbs + k);
    rec[0] = f[0];
    rec[1] = f[1];
    rec[2] = f[2];
    rec[3] = bp;
    rec[4] = cosw;
    rec[5] = mint;
    if (ADJ) rec[6] = dfr;
  }
}

// ------------------------------------------------------------------ photon final gather
// PhotonLTEIntegrator::_FinalGather (PhotonLTEIntegrator.cpp:387-577) as a wavefront over a pass's
// shading points: lookup inputs -> 32 nearest indirect photons (gather kernel) -> BSDF and cone
// samples with their gather rays -> closest hit -> hit scattering and irradiance lookup inputs ->
// nearest irradiance photon (gather kernel) -> per-sample ordered sum.

// Counter-based stream of the final gather's stratified samples (dims up to 4096 per sample).
__device__ __forceinline__ double fg_rnd(uint64_t seed, uint64_t sample, uint32_t dim) {
  uint64_t h = mix64((seed ^ 0x6A09E667F3BCC909ull) ^ mix64(sample * 4096ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}

constexpr double kCosGather = 0.9848;  // PhotonLTEIntegrator.cpp:396 (10 degrees)

// Gather rays are stored sample-major within blocks of kFgBlock slots: sample j (0 .. 2 gs: BSDF
// samples, then cone samples) of slot i is ray (i / B) B 2gs + j B + i % B. A wave of consecutive
// rays then holds one sample stratum of 64 neighbouring slots, and the per-slot reads of the
// resolve are consecutive across a wave's lanes. Slots past the batch (up to the block) hold rays
// that cast nothing.
__device__ __forceinline__ uint64_t fg_ray(uint64_t i, uint32_t j, uint32_t gs) {
  return (i / kFgBlock) * (kFgBlock * 2ull * gs) + (uint64_t)j * kFgBlock + (i % kFgBlock);
}

__global__ __launch_bounds__(256) void k_fg_lookup_in(FgArgs f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.a.batch) return;
  Point pt;
  const bool hit = load_point(f.a, i, pt);
  const double nan = __builtin_nan("");
  const D3 p = hit ? pt.p : D3{nan, nan, nan};
  const D3 n = hit ? pt.g.ng : D3{0.0, 0.0, 1.0};
  f.s_pts[3 * i] = p.x;
  f.s_pts[3 * i + 1] = p.y;
  f.s_pts[3 * i + 2] = p.z;
  f.s_nrm[3 * i] = n.x;
  f.s_nrm[3 * i + 1] = n.y;
  f.s_nrm[3 * i + 2] = n.z;
}

// The found photons' directions of a block's kFgBlock slots, staged in LDS once per block and read
// by every sample of the slot: entry (j, c) of slot lane is dirs[(3 j + c) kFgBlock + lane], so a
// wave's lanes (consecutive slots, one j) read consecutive doubles.
constexpr uint32_t kFgSampleThreads = 256;
struct FgDirs {
  double v[32 * 3 * kFgBlock];
};

__device__ __forceinline__ D3 photon_dir(const FgDirs &d, uint32_t lane, uint32_t j) {
  const double *p = d.v + 3 * j * kFgBlock + lane;
  return D3{p[0], p[kFgBlock], p[2 * kFgBlock]};
}

// The photon-sampling pdf of direction ex: a 10-degree cone around each found photon's direction.
__device__ __forceinline__ double photon_pdf(const FgDirs &d, uint32_t lane, uint32_t found, D3 ex) {
  const double cone_pdf = 1.0 / (2.0 * kPi * (1.0 - kCosGather));  // UniformConePDF
  double pp = 0.0;
  for (uint32_t j = 0; j < found; ++j)
    if (dot3(photon_dir(d, lane, j), ex) > kCosGather) pp += cone_pdf;
  return pp * (found > 0 ? 1.0 / found : 0.0);
}

// The BSDF sample k (s < gs, :470-500) or the cone sample k = s - gs around a found photon's
// direction (s >= gs, :502-532) of a shaded slot; returns whether it casts a gather ray.
__device__ __forceinline__ bool fg_sample(const FgArgs &f, uint64_t i, uint32_t s, const Point &pt, const Bsdf &b,
                                          uint32_t found, const FgDirs &d, uint32_t lane, D3 &ex, double wv[3]) {
  const DlArgs &a = f.a;
  const uint32_t gs = f.gs, gq = f.gs_sqrt;
  const bool cone = s >= gs;
  const uint32_t k = cone ? s - gs : s;
  const uint64_t sidx = pixel_of(a, i) * a.spp + pass_of(a, i);
  const double inv_n = 1.0 / gs, inv_q = 1.0 / gq;
  const uint32_t x = k % gq, y = k / gq;
  if (!cone) {
    // StratifiedSampling1D / 2D with jitter (SamplingRoutines.h:327-373): dims [0, gs), [gs, 3 gs)
    const double uc = (k + fg_rnd(a.seed, sidx, k)) * inv_n;
    const double u0 = (x + fg_rnd(a.seed, sidx, gs + 2 * k)) * inv_q;
    const double u1 = (y + fg_rnd(a.seed, sidx, gs + 2 * k + 1)) * inv_q;
    double bp, fv[3];
    D3 wo_l, wi_l;
    b.sample(pt.wo, u0, u1, uc, ex, bp, fv, wo_l, wi_l);
    if (!(bp > 0.0 && !(fv[0] == 0.0 && fv[1] == 0.0 && fv[2] == 0.0))) return false;
    const double ppdf = photon_pdf(d, lane, found, ex);
    const double w = power_heuristic((double)gs, bp, (double)gs, ppdf);
    const double kk = w * fabs(dot3(ex, b.n)) / bp;
    for (int c = 0; c < 3; ++c) wv[c] = fv[c] * kk;
    return true;
  }
  if (found == 0) return false;
  // dims [3 gs, 4 gs) and [4 gs, 6 gs)
  const double us = (k + fg_rnd(a.seed, sidx, 3 * gs + k)) * inv_n;
  const double u0 = (x + fg_rnd(a.seed, sidx, 4 * gs + 2 * k)) * inv_q;
  const double u1 = (y + fg_rnd(a.seed, sidx, 4 * gs + 2 * k + 1)) * inv_q;
  const uint32_t sj = (uint32_t)(us * found);
  const D3 e1 = photon_dir(d, lane, sj);
  // MathRoutines::CoordinateSystem (MathRoutines.h:298-313)
  D3 e2;
  if (fabs(e1.x) > fabs(e1.y)) {
    const double inv = 1.0 / sqrt(e1.x * e1.x + e1.z * e1.z);
    e2 = D3{-e1.z * inv, 0.0, e1.x * inv};
  } else {
    const double inv = 1.0 / sqrt(e1.y * e1.y + e1.z * e1.z);
    e2 = D3{0.0, e1.z * inv, -e1.y * inv};
  }
  const D3 e3 = cross3(e1, e2);
  // SamplingRoutines::UniformConeSampling (SamplingRoutines.h:300-310)
  const double ct = (1.0 - u0) + u0 * kCosGather;
  const double st = sqrt(1.0 - ct * ct);
  const double phi = u1 * 2.0 * kPi;
  const double lx = st * cos(phi), ly = st * sin(phi);
  ex = add3(add3(mul3(e2, lx), mul3(e3, ly)), mul3(e1, ct));
  double fv[3];
  b.evaluate(pt.wo, ex, fv);
  if (fv[0] == 0.0 && fv[1] == 0.0 && fv[2] == 0.0) return false;
  const double bp = b.pdf(pt.wo, ex);
  const double ppdf = photon_pdf(d, lane, found, ex);
  const double w = power_heuristic((double)gs, ppdf, (double)gs, bp);
  const double kk = w * fabs(dot3(ex, b.n)) / ppdf;
  for (int c = 0; c < 3; ++c) wv[c] = fv[c] * kk;
  return true;
}

// One block per kFgBlock slots, one lane per slot: the slot's point, BSDF and found photon
// directions are loaded once, then the block's waves take the 2 gs samples in turn (wave w takes
// samples w, w + 4, ...) and write their gather rays.
__global__ __launch_bounds__(kFgSampleThreads) void k_fg_samples(FgArgs f) {
  __shared__ FgDirs dirs;
  __shared__ uint32_t wave_rays[kFgSampleThreads / 64];
  const DlArgs &a = f.a;
  const uint32_t gs = f.gs, lane = threadIdx.x % kFgBlock, wave = threadIdx.x / kFgBlock;
  const uint64_t i = (uint64_t)blockIdx.x * kFgBlock + lane;
  Point pt;
  const bool shade = i < a.batch && load_point(a, i, pt);
  const uint32_t found = (shade && f.near_count) ? f.near_count[i] : 0u;
  for (uint32_t j = wave; j < found; j += kFgSampleThreads / kFgBlock) {
    const double *dv = f.cdir + 3 * (size_t)f.near_ids[i * 32 + j];
    for (int c = 0; c < 3; ++c) dirs.v[(3 * j + c) * kFgBlock + lane] = dv[c];
  }
  __syncthreads();
  Bsdf b;
  if (shade) point_bsdf(a, i, pt.prim, b);
  uint32_t rays = 0;
  for (uint32_t s = wave; s < 2 * gs; s += kFgSampleThreads / kFgBlock) {
    D3 ex{0, 0, 0};
    double wv[3] = {0.0, 0.0, 0.0}, mint = 0.0;
    const bool want = shade && fg_sample(f, i, s, pt, b, found, dirs, lane, ex, wv);
    if (want) mint = next_min_t(a.geom.prims[pt.prim], pt.g, ex);
    rays += (uint32_t)__popcll(__ballot(want));
    // A sample that casts no ray gets a ray that starts outside the scene box, points away from
    // it and has tmax < tmin: its root test culls it, its record is a miss, its weight is 0.
    const uint64_t g = fg_ray(i, s, gs);
    if (want)
      put_ray(f.rays + g, pt.p, ex, mint, __builtin_huge_val());
    else
      put_ray(f.rays + g, D3{f.hole_o[0], f.hole_o[1], f.hole_o[2]}, D3{f.hole_d[0], f.hole_d[1], f.hole_d[2]}, 0.0,
              -1.0);
    f.w[3 * g] = wv[0];
    f.w[3 * g + 1] = wv[1];
    f.w[3 * g + 2] = wv[2];
  }
  // the gather rays cast, one atomic per block
  if (lane == 0) wave_rays[wave] = rays;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < kFgSampleThreads / kFgBlock; ++w) total += wave_rays[w];
    if (total) atomicAdd(reinterpret_cast<unsigned long long *>(f.count), (unsigned long long)total);
  }
}

// BSDF::TotalScattering(incident, 5x5 stratified samples, BSDF_ALL_REFLECTION) of a matte /
// plastic surface (BSDF.cpp:209-224): Lambertian::TotalScattering = R (Lambertian.cpp:32-36) plus,
// for plastic, BxDF::TotalScattering of the microfacet lobe (BxDF.cpp:43-68) through
// Microfacet::Sample (Microfacet.h:112-128) and BlinnDistribution::Sample (BlinnDistribution.h:
// 81-98). BSDF::TotalScattering hands the lobes the world-space incident vector, as here.
// Quadrature point (x, y) reflects the incident vector i about the half vector h (cos_h = ct_x,
// azimuth phi_y, z flipped to i's side), so the half vector Microfacet::Evaluate rebuilds from the
// pair is h itself up to sign: cos_h = ct_x and |i.h'| = |ex.h'| = |c| with c = i.h. The term
// f |cos_o| / pdf then has D = (e + 2) / (2 pi) ct^e over pdf = (e + 1) ct^e / (8 pi |c|), whose
// ct^e cancels (ct >= 0.1^(1 / (e + 1)), so it never underflows), leaving
// ks Fr(|c|) G |c| (e + 2) / ((e + 1) |cos_i|): the same sum, term by term, up to rounding, with
// five pow and five sincos per hit instead of a hundred transcendentals.
// The quadrature's angles depend only on the sample index (azimuth) and the material's exponent
// (cos_h), so k_fg_quad computes them once per call with the same expressions, and
// total_scattering reads them: q = the azimuth table, qm = the material's cos_h / sin_h.
__device__ __forceinline__ void quad_azimuths(double *q) {
  for (int x = 0; x < 5; ++x) {
    const double u = (x + 0.5) * (1.0 / 5);
    sincos(u * 2.0 * kPi, &q[5 + x], &q[x]);
  }
}
__device__ __forceinline__ void quad_heights(double e, double *qm) {
  for (int x = 0; x < 5; ++x) {
    const double u = (x + 0.5) * (1.0 / 5);
    qm[x] = pow(u, 1.0 / (e + 1.0));
    qm[5 + x] = sqrt(dmax(0.0, 1.0 - qm[x] * qm[x]));
  }
}

__device__ __forceinline__ void total_scattering(const Bsdf &b, D3 incident, const double *__restrict__ q,
                                                 const double *__restrict__ qm, double ts[3]) {
  double r[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 3; ++c) r[c] = r[c] + b.kd[c];
  if (b.plastic) {
    double ct[5], st[5], cp[5], sp[5];
    for (int x = 0; x < 5; ++x) {
      cp[x] = q[x];
      sp[x] = q[5 + x];
      ct[x] = qm[x];
      st[x] = qm[5 + x];
    }
    const double ci = fabs(incident.z);
    const double kfac = (b.e + 2.0) / (b.e + 1.0);
    double m[3] = {0.0, 0.0, 0.0};
    for (int y = 0; y < 5; ++y)
      for (int x = 0; x < 5; ++x) {
        D3 h{st[x] * cp[y], st[x] * sp[y], ct[x]};
        if (incident.z * h.z < 0.0) h.z *= -1.0;
        const double c = dot3(incident, h);
        const D3 ex = sub3(mul3(h, 2.0 * c), incident);
        if (incident.z * ex.z < 0.0) continue;  // pdf 0
        const double co = fabs(ex.z);
        if (ci < kEps || co < kEps) continue;  // Microfacet::Evaluate returns black
        const D3 h2 = add3(incident, ex);
        if (fabs(h2.x) < kEps && fabs(h2.y) < kEps && fabs(h2.z) < kEps) continue;
        const double ac = fabs(c);
        const double tmp = 2.0 * ct[x] / ac;
        const double g = dmin(1.0, dmin(ci * tmp, co * tmp));
        const double w = fresnel_dielectric(ac) * g * ac * kfac / ci;
        for (int k = 0; k < 3; ++k) m[k] += w * b.ks[k];
      }
    for (int c = 0; c < 3; ++c) {
      m[c] = m[c] * (1.0 / 25.0);
      m[c] = dmin(1.0, dmax(0.0, m[c]));
      r[c] = r[c] + m[c];
    }
  }
  for (int c = 0; c < 3; ++c) ts[c] = dmin(1.0, dmax(0.0, r[c]));
}

// One thread per gather ray: the hit's geometry, total reflection scattering toward the gather
// direction, and the irradiance lookup's point and geometric normal (:545-570).
// The total scattering of a gather hit toward the gather direction (k_fg_hit for matte surfaces,
// k_fg_plastic for plastic ones).
__device__ __forceinline__ void gather_scattering(const FgArgs &f, uint64_t g, uint32_t prim, double ts[3]) {
  const skw_ray &r = f.rays[g];
  const D3 gd{-r.d[0], -r.d[1], -r.d[2]};
  Bsdf b2;
  const DevMat &m = f.a.mats[prim];
  b2.plastic = m.type == 1;
  for (int c = 0; c < 3; ++c) {
    b2.kd[c] = m.kd[c];
    b2.ks[c] = m.ks[c];
  }
  b2.e = m.exponent;
  total_scattering(b2, gd, f.quad, f.quad + 10 + 10 * (size_t)prim, ts);
}

// Plastic gather hits are listed (one atomic per block) and their 25-point quadrature runs in
// k_fg_plastic with one listed hit per lane: in k_fg_hit the quadrature would run with only the
// plastic lanes of each wave active.
__global__ __launch_bounds__(256) void k_fg_hit(FgArgs f) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = g < f.slots_padded * 2ull * f.gs;
  const DlArgs &a = f.a;
  const double nan = __builtin_nan("");
  D3 qp{nan, nan, nan}, qn{0.0, 0.0, 1.0};
  double ts[3] = {0.0, 0.0, 0.0};
  uint8_t side = 0;
  bool plastic = false;
  // Misses (samples that cast no ray included) get a NaN lookup point: the irradiance lookup runs
  // over every slot and a NaN point descends one path and finds nothing.
  skw_hit h;
  h.prim = kMiss;
  if (in) h = f.hits[g];
  if (h.prim != kMiss) {
    const skw_ray &r = f.rays[g];
    const D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
    const HitGeom g2 = hit_geometry(a.geom, o, d, h.prim, h.tri);
    const D3 gd = mul3(d, -1.0);
    plastic = a.mats[h.prim].type == 1;
    if (!plastic) gather_scattering(f, g, h.prim, ts);
    side = dot3(gd, g2.ng) > 0.0 ? 1 : 0;
    qp = g2.p;
    qn = g2.ng;
  }
  const int32_t at = append_slot_block(f.plastic_count, plastic);
  if (plastic) f.plastic[at] = (uint32_t)g;
  if (!in) return;
  if (!plastic)
    for (int c = 0; c < 3; ++c) f.ts[3 * g + c] = ts[c];
  f.q_pts[3 * g] = qp.x;
  f.q_pts[3 * g + 1] = qp.y;
  f.q_pts[3 * g + 2] = qp.z;
  f.q_nrm[3 * g] = qn.x;
  f.q_nrm[3 * g + 1] = qn.y;
  f.q_nrm[3 * g + 2] = qn.z;
  f.side[g] = side;
  if (f.keys) {
    // 2^B cells per axis over the scene box, Morton-interleaved into the key's top 3B bits;
    // misses sort last
    uint32_t key = 0xFFFFFFFFu;
    if (h.prim != kMiss) {
      const uint32_t B = f.morton_bits, lo_bit = 32u - 3u * B;
      const double top = (double)((1u << B) - 1u);
      const double c[3] = {(qp.x - f.lo[0]) * f.cell[0], (qp.y - f.lo[1]) * f.cell[1], (qp.z - f.lo[2]) * f.cell[2]};
      key = 0;
      for (int ax = 0; ax < 3; ++ax) {
        const uint32_t v = (uint32_t)dmin(top, dmax(0.0, c[ax]));
        for (uint32_t bit = 0; bit < B; ++bit) key |= ((v >> bit) & 1u) << (lo_bit + 3 * bit + ax);
      }
    }
    f.keys[g] = key;
    f.order[g] = (uint32_t)g;
  }
}

// The total scattering of the listed plastic gather hits (a grid-stride loop over the list).
__global__ __launch_bounds__(256) void k_fg_plastic(FgArgs f) {
  const uint64_t n = *f.plastic_count;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t g = f.plastic[t];
    double ts[3];
    gather_scattering(f, g, f.hits[g].prim, ts);
    for (int c = 0; c < 3; ++c) f.ts[3 * g + c] = ts[c];
  }
}

// The quadrature table of total_scattering: one thread per material (plastic ones), thread
// n_mats the azimuths.
__global__ __launch_bounds__(256) void k_fg_quad(FgArgs f) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == f.n_mats) quad_azimuths(f.quad);
  if (i >= f.n_mats) return;
  const DevMat &m = f.a.mats[i];
  if (m.type == 1) quad_heights(m.exponent, f.quad + 10 + 10 * (size_t)i);
}

// One thread per slot: the sample's gather rays in the reference's order (BSDF samples, then
// cone samples), each adding TotalScattering x irradiance / pi x weight (:561-573); / gather
// samples (Spectrum::operator/ multiplies by the reciprocal). Film and per-sample output.
__global__ __launch_bounds__(256) void k_fg_resolve(FgArgs f) {
  const DlArgs &a = f.a;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    a.totals[0] += a.batch;
    a.totals[1] += *f.count;
  }
  if (i >= a.batch) return;
  const uint32_t gs = f.gs;
  double rad[3] = {0.0, 0.0, 0.0};
  for (uint32_t j = 0; j < 2 * gs; ++j) {
    const uint64_t r = fg_ray(i, j, gs);
    // A gather ray that missed (or a sample that cast none) has a NaN lookup point (k_fg_hit),
    // which finds no irradiance photon: its id is kMiss.
    const uint32_t id = f.irr_id[r];
    if (id == kMiss) continue;
    const bool ext = f.side[r] != 0;
    const float *es = ext ? f.e_ext + 3 * (size_t)id : f.e_int + 3 * (size_t)id;
    const float *eo = ext ? f.e_int + 3 * (size_t)id : f.e_ext + 3 * (size_t)id;
    for (int c = 0; c < 3; ++c) {
      double t = 0.0;
      t = t + f.ts[3 * (size_t)r + c] * (double)es[c];
      t = t + 0.0 * (double)eo[c];  // no transmission lobe: its total scattering is 0
      t = t * kInvPi;
      rad[c] = rad[c] + (t * 1.0) * f.w[3 * (size_t)r + c];
    }
  }
  const double inv = gs > 0 ? 1.0 / (double)gs : 0.0;
  const uint64_t pix = pixel_of(a, i);
  for (int c = 0; c < 3; ++c) {
    const double L = rad[c] * inv;
    // several passes per launch: the slot's value, added to the film in pass order by k_dl_accum
    if (a.slot_L) a.slot_L[4 * i + c] = L;
    else a.film[3 * pix + c] += L;
    if (f.radiance) f.radiance[3 * (pix * a.spp + pass_of(a, i)) + c] = L;
  }
  if (a.slot_L) a.slot_L[4 * i + 3] = 0.0;  // the final gather drops no sample
}

// Per-sample sum and film accumulation. ADJ: instead of the film, accumulates the sample's
// contribution to d_grad: for L = Le + (1/nls) sum_k f_k * le_k * W_k + (1/nbs) sum_k f_k * le_k * W_k
// with the weights W fixed, dL/d(emission l) = base radiance of l times the terms that use l,
// dL/d(reflectance g) = (dkd / pi) * le * W summed, dL/d(roughness g) = ks * dfr * le * W summed,
// where g is the group of the primitive the camera ray hit. The sample's dLoss/dL is
// adj[pixel] / spp. Sums go to LDS (ds_add_f64), then one global FP64 atomic per parameter per block.
// MODE 0: film. MODE 1 (ADJ): the contraction above. MODE 2 (TAPE): film, and the sample's
// derivatives themselves (dL[c] per reflectance, roughness and each emission scale) into its
// tape row, for a later contraction with any dLoss/dImage.
template <int MODE>
__global__ __launch_bounds__(256) void k_dl_resolve(DlArgs a) {
  constexpr bool ADJ = MODE == 1, TAPE = MODE == 2;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    a.totals[0] += a.batch;
    a.totals[1] += a.counts[0];
    a.totals[2] += a.counts[1];
  }
  __shared__ double gsum[ADJ ? kDlMaxParams : 1];
  const uint32_t G = a.n_groups, n_params = 2 * a.n_groups + a.n_lights;
  if (ADJ) {
    for (uint32_t t = threadIdx.x; t < n_params; t += blockDim.x) gsum[t] = 0.0;
    __syncthreads();
  }
  const bool in = i < a.batch;
  skw_hit h;
  h.prim = kMiss;
  if (in) h = a.prim_hits[i];
  double L[3] = {0.0, 0.0, 0.0};
  if (h.prim != kMiss) {
    for (int c = 0; c < 3; ++c) L[c] = a.le[3 * i + c];
    double cdl[3] = {0.0, 0.0, 0.0};
    const uint32_t nl = a.n_lights;
    if (nl > 0) {
      double dl[3] = {0.0, 0.0, 0.0};
      for (uint32_t k = 0; k < a.nls; ++k) {
        int32_t j = a.shadow_idx[i * a.nls + k];
        if (j < 0 || a.occluded[j]) continue;
        const double *c3 = a.shadow_c + 3 * (i * a.nls + k);
        for (int c = 0; c < 3; ++c) dl[c] += c3[c];
      }
      for (int c = 0; c < 3; ++c) cdl[c] = dl[c] * (1.0 / (double)a.nls);
      double db[3] = {0.0, 0.0, 0.0};
      for (uint32_t k = 0; k < a.nbs; ++k) {
        int32_t j = a.bsdf_idx[i * a.nbs + k];
        if (j < 0) continue;
        double *rec = a.bsdf_rec + 8 * (i * a.nbs + k);
        if (ADJ || TAPE) rec[7] = -1.0;  // no light reached
        const skw_hit h2 = a.bsdf_hits[j];
        if (h2.prim == kMiss) continue;
        const DevMat &m2 = a.mats[h2.prim];
        if (m2.light == kMiss) continue;
        const skw_ray &r2 = a.bsdf_rays[j];
        D3 wi{r2.d[0], r2.d[1], r2.d[2]};
        // DiffuseAreaLightSource::Radiance tests the DG geometric normal of the hit.
        HitGeom g2 = hit_geometry(a.geom, D3{r2.o[0], r2.o[1], r2.o[2]}, wi, h2.prim, h2.tri);
        if (!(dot3(mul3(wi, -1.0), g2.ng) > 0.0)) continue;
        const DevLight &Lt = a.lights[m2.light];
        if (Lt.radiance[0] == 0.0 && Lt.radiance[1] == 0.0 && Lt.radiance[2] == 0.0) continue;
        const DevLightTri &T = a.ltris[Lt.tri_off + h2.tri];
        double dist = h2.t - rec[5];
        double ac = fabs(dot3(wi, D3{T.n[0], T.n[1], T.n[2]}));
        double lpdf = ac > kEps ? (dist * dist) / (Lt.area * ac) : 0.0;
        const double *cdf = a.cdf + i * nl;
        double cp = m2.light == 0 ? cdf[0] : cdf[m2.light] - cdf[m2.light - 1];
        double w = power_heuristic((double)a.nbs, rec[3], (double)a.nls, lpdf * cp);
        w *= rec[4];
        const double *re = a.rad_e + 3 * m2.light;
        for (int c = 0; c < 3; ++c) db[c] += rec[c] * re[c] * w;
        if (ADJ || TAPE) rec[7] = w;
      }
      for (int c = 0; c < 3; ++c) cdl[c] += db[c] * (1.0 / (double)a.nbs);
    }
    for (int c = 0; c < 3; ++c) L[c] = L[c] + cdl[c];
  }
  double lum = 0.0;
  for (int c = 0; c < 3; ++c) lum += kLumW[c] * L[c];
  bool bad = isnan(L[0]) || isnan(L[1]) || isnan(L[2]) || lum < -kEps || isinf(L[0]) || isinf(L[1]) || isinf(L[2]);
  if (!ADJ) {
    if (a.sample_L) {
      // the sum before the renderer's rule: the composition adds the other terms first
      if (in) {
        double *o = a.sample_L + 3 * (pixel_of(a, i) * a.spp + pass_of(a, i));
        for (int c = 0; c < 3; ++c) o[c] = L[c];
      }
    } else if (a.slot_L) {
      // several passes per launch: k_dl_accum adds the launch's samples of a pixel in pass order
      if (in) {
        double *o = a.slot_L + 4 * i;
        for (int c = 0; c < 3; ++c) o[c] = L[c];
        o[3] = bad ? 1.0 : 0.0;
      }
    } else if (in && !bad) {
      double *f = a.film + 3 * pixel_of(a, i);
      for (int c = 0; c < 3; ++c) f[c] += L[c];
    }
    if (TAPE && in) {
      // row: [0..2] reflectance, [3..5] roughness, [6 + 3l + c] emission scale of light l
      double *J = a.tape_j + tape_row(a, i) * (uint64_t)a.tape_stride;
      for (uint32_t t = 0; t < a.tape_stride; ++t) J[t] = 0.0;
      uint32_t grp = kMiss;
      if (h.prim != kMiss && !bad) {
        const DevMat &me = a.mats_e[h.prim];
        grp = me.group;
        double jr[3] = {0.0, 0.0, 0.0}, js[3] = {0.0, 0.0, 0.0};
        const int32_t cl = a.cam_light[i];
        if (cl >= 0)
          for (int c = 0; c < 3; ++c) J[6 + 3 * cl + c] += a.rad_base[3 * cl + c];
        if (a.n_lights > 0) {
          const double inv_nls = 1.0 / (double)a.nls, inv_nbs = 1.0 / (double)a.nbs;
          for (uint32_t k = 0; k < a.nls; ++k) {
            int32_t j = a.shadow_idx[i * a.nls + k];
            if (j < 0 || a.occluded[j]) continue;
            const double *x = a.shadow_aux + 6 * (i * a.nls + k);
            const uint32_t l = (uint32_t)x[5];
            const double W = x[3] * inv_nls;
            const double *re = a.rad_e + 3 * l, *rb = a.rad_base + 3 * l;
            for (int c = 0; c < 3; ++c) {
              J[6 + 3 * l + c] += x[c] * rb[c] * W;
              jr[c] += re[c] * W;
              js[c] += re[c] * me.ks[c] * x[4] * W;
            }
          }
          for (uint32_t k = 0; k < a.nbs; ++k) {
            int32_t j = a.bsdf_idx[i * a.nbs + k];
            if (j < 0) continue;
            const double *rec = a.bsdf_rec + 8 * (i * a.nbs + k);
            if (rec[7] < 0.0) continue;
            const uint32_t l = a.mats[a.bsdf_hits[j].prim].light;
            const double W = rec[7] * inv_nbs;
            const double *re = a.rad_e + 3 * l, *rb = a.rad_base + 3 * l;
            for (int c = 0; c < 3; ++c) {
              J[6 + 3 * l + c] += rec[c] * rb[c] * W;
              jr[c] += re[c] * W;
              js[c] += re[c] * me.ks[c] * rec[6] * W;
            }
          }
        }
        for (int c = 0; c < 3; ++c) {
          J[c] = jr[c] * me.dkd * kInvPi;
          J[3 + c] = js[c];
        }
      }
      a.tape_group[tape_row(a, i)] = grp;
    }
    return;
  }
  if (in && h.prim != kMiss && !bad) {
    const uint64_t pix = pixel_of(a, i);
    const double inv_spp = 1.0 / (double)a.spp;
    double wa[3];
    for (int c = 0; c < 3; ++c) wa[c] = a.adj[3 * pix + c] * inv_spp;
    const DevMat &me = a.mats_e[h.prim];
    double g_refl = 0.0, g_rough = 0.0;
    const int32_t cl = a.cam_light[i];
    if (cl >= 0) {
      const double *rb = a.rad_base + 3 * cl;
      atomicAdd(&gsum[2 * G + cl], wa[0] * rb[0] + wa[1] * rb[1] + wa[2] * rb[2]);
    }
    if (a.n_lights > 0) {
      const double inv_nls = 1.0 / (double)a.nls, inv_nbs = 1.0 / (double)a.nbs;
      for (uint32_t k = 0; k < a.nls; ++k) {
        int32_t j = a.shadow_idx[i * a.nls + k];
        if (j < 0 || a.occluded[j]) continue;
        const double *x = a.shadow_aux + 6 * (i * a.nls + k);
        const uint32_t l = (uint32_t)x[5];
        const double W = x[3] * inv_nls;
        const double *re = a.rad_e + 3 * l, *rb = a.rad_base + 3 * l;
        double ge = 0.0, gr = 0.0, gs = 0.0;
        for (int c = 0; c < 3; ++c) {
          ge += wa[c] * x[c] * rb[c];
          gr += wa[c] * re[c];
          gs += wa[c] * re[c] * me.ks[c];
        }
        atomicAdd(&gsum[2 * G + l], ge * W);
        g_refl += gr * W;
        g_rough += gs * x[4] * W;
      }
      for (uint32_t k = 0; k < a.nbs; ++k) {
        int32_t j = a.bsdf_idx[i * a.nbs + k];
        if (j < 0) continue;
        const double *rec = a.bsdf_rec + 8 * (i * a.nbs + k);
        if (rec[7] < 0.0) continue;
        const uint32_t l = a.mats[a.bsdf_hits[j].prim].light;
        const double W = rec[7] * inv_nbs;
        const double *re = a.rad_e + 3 * l, *rb = a.rad_base + 3 * l;
        double ge = 0.0, gr = 0.0, gs = 0.0;
        for (int c = 0; c < 3; ++c) {
          ge += wa[c] * rec[c] * rb[c];
          gr += wa[c] * re[c];
          gs += wa[c] * re[c] * me.ks[c];
        }
        atomicAdd(&gsum[2 * G + l], ge * W);
        g_refl += gr * W;
        g_rough += gs * rec[6] * W;
      }
    }
    if (me.group != kMiss) {
      atomicAdd(&gsum[me.group], g_refl * me.dkd * kInvPi);
      atomicAdd(&gsum[G + me.group], g_rough);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n_params; t += blockDim.x) {
    const double v = gsum[t];
    if (v != 0.0) unsafeAtomicAdd(a.grad + t, v);
  }
}

// Contraction of a tape with dLoss/dImage: grad[j] += sum over samples and channels of
// adj[pixel] / spp * row[j-part]; grid-stride, LDS sums, one global atomic per parameter per block.
__global__ __launch_bounds__(256) void k_dl_tape_adjoint(DlTapeArgs t) {
  __shared__ double gsum[kDlMaxParams];
  const uint32_t G = t.n_groups, nl = t.n_lights, n_params = 2 * G + nl;
  for (uint32_t p = threadIdx.x; p < n_params; p += blockDim.x) gsum[p] = 0.0;
  __syncthreads();
  const double inv_spp = 1.0 / (double)t.spp;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < t.n_samples;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t pix = pixel_of(t.width, t.height, s % t.batch);
    double wa[3];
    for (int c = 0; c < 3; ++c) wa[c] = t.adj[3 * pix + c] * inv_spp;
    const double *J = t.tape_j + s * (uint64_t)t.stride;
    const uint32_t grp = t.tape_group[s];
    if (grp != kMiss) {
      atomicAdd(&gsum[grp], wa[0] * J[0] + wa[1] * J[1] + wa[2] * J[2]);
      atomicAdd(&gsum[G + grp], wa[0] * J[3] + wa[1] * J[4] + wa[2] * J[5]);
    }
    for (uint32_t l = 0; l < nl; ++l) {
      const double *e = J + 6 + 3 * l;
      const double v = wa[0] * e[0] + wa[1] * e[1] + wa[2] * e[2];
      if (v != 0.0) atomicAdd(&gsum[2 * G + l], v);
    }
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < n_params; p += blockDim.x) {
    const double v = gsum[p];
    if (v != 0.0) unsafeAtomicAdd(t.grad + p, v);
  }
}

// Clears the film and the ray totals (a kernel rather than a memset, so a captured frame replays
// it faithfully).
__global__ __launch_bounds__(256) void k_dl_clear(DlArgs a) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)a.width * a.height * 3;
  if (i < 9) a.totals[i] = 0;  // ray totals and the traversal test sums of a counting frame
  if (i < n) a.film[i] = 0.0;
}

// Several passes per launch: the launch's samples of each pixel added to the film in pass order,
// as the one-pass launches add them (dropped samples skipped), one thread per pixel slot.
__global__ __launch_bounds__(256) void k_dl_accum(DlArgs a) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (a.batch >> a.ppl_log2)) return;
  const uint64_t first = q << a.ppl_log2;
  double *f = a.film + 3 * pixel_of(a, first);
  double acc[3] = {f[0], f[1], f[2]};
  for (uint64_t k = 0; k < (1ull << a.ppl_log2); ++k) {
    const double *o = a.slot_L + 4 * (first + k);
    if (o[3] == 0.0)
      for (int c = 0; c < 3; ++c) acc[c] += o[c];
  }
  for (int c = 0; c < 3; ++c) f[c] = acc[c];
}

__global__ __launch_bounds__(256) void k_dl_finish(DlArgs a, float *rgb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (a.batch >> a.ppl_log2)) return;
  const uint64_t pix = a.pix0 + i;
  for (int c = 0; c < 3; ++c) rgb[3 * pix + c] = (float)(a.film[3 * pix + c] / (double)a.spp);
}

// PhotonLTEIntegrator::_SurfaceRadiance (PhotonLTEIntegrator.cpp:320-385) per sample: ((Le + direct)
// + caustic) + final gather, where the caustic lookup of a photon-less caustic map is a zero
// spectrum (matte and plastic surfaces scatter nothing specularly); a missed camera ray has 0 in
// both inputs. Then the renderer's rule (SamplerBasedRenderer.cpp:458-478: NaN, negative
// luminance or infinite samples are dropped), the passes summed in order, and the pixel mean.
__global__ __launch_bounds__(256) void k_pm_combine(const double *sample_dl, const double *sample_fg, uint64_t npix,
                                                    uint32_t spp, float *rgb) {
  const uint64_t pix = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= npix) return;
  double acc[3] = {0.0, 0.0, 0.0};
  for (uint32_t s = 0; s < spp; ++s) {
    const uint64_t k = 3 * (pix * spp + s);
    double L[3];
    for (int c = 0; c < 3; ++c) L[c] = (sample_dl[k + c] + 0.0) + (sample_fg ? sample_fg[k + c] : 0.0);
    double lum = 0.0;
    for (int c = 0; c < 3; ++c) lum += kLumW[c] * L[c];
    const bool bad = isnan(L[0]) || isnan(L[1]) || isnan(L[2]) || lum < -kEps || isinf(L[0]) || isinf(L[1]) ||
                     isinf(L[2]);
    if (!bad)
      for (int c = 0; c < 3; ++c) acc[c] += L[c];
  }
  for (int c = 0; c < 3; ++c) rgb[3 * pix + c] = (float)(acc[c] / (double)spp);
}

__global__ __launch_bounds__(256) void k_dl_finish_f64(DlArgs a, double *img) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (a.batch >> a.ppl_log2)) return;
  const uint64_t pix = a.pix0 + i;
  for (int c = 0; c < 3; ++c) img[3 * pix + c] = a.film[3 * pix + c] / (double)a.spp;
}

__global__ __launch_bounds__(256) void k_dl_clear_grad(double *grad, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grad[i] = 0.0;
}

// Material of a grouped primitive at one parameter vector: gray kd = clamp(theta[g], 0, 1) and
// Blinn exponent 1/clamp(theta[G+g], 0.001, 1) clamped to [0, 1000], as the scene's own materials
// (MatteMaterial.cpp:33-34, PlasticMaterial.cpp:37-45); derivatives are 0 where a clamp holds.
__device__ void param_material(const DevMat &base, const double *theta, uint32_t G, DevMat &o) {
  o = base;
  if (base.group == kMiss) return;
  const double rho = theta[base.group];
  const double kd = rho < 0.0 ? 0.0 : (rho > 1.0 ? 1.0 : rho);
  for (int c = 0; c < 3; ++c) o.kd[c] = kd;
  o.dkd = (rho > 0.0 && rho < 1.0) ? 1.0 : 0.0;
  const double r = theta[G + base.group];
  const double rc = r < 0.001 ? 0.001 : (r > 1.0 ? 1.0 : r);
  const double e = 1.0 / rc;
  o.exponent = e < 0.0 ? 0.0 : (e > 1000.0 ? 1000.0 : e);
  o.de_dr = (r > 0.001 && r < 1.0) ? -1.0 / (r * r) : 0.0;
}

// Sampling and evaluation tables of one call. Light intensity as the scene's light table has it:
// luminance of (pi * area) * radiance over pi times the light's bounding-box area
// (IrradianceLightsSamplingStrategy.cpp:19-31).
__global__ __launch_bounds__(256) void k_dl_params(DlParamArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n_prims) {
    const DevMat base = a.base_mats[i];
    DevMat m;
    param_material(base, a.theta_s, a.n_groups, m);
    a.mats_s[i] = m;
    if (a.mats_e != a.mats_s) {
      param_material(base, a.theta_e, a.n_groups, m);
      a.mats_e[i] = m;
    }
  }
  if (i < a.n_lights) {
    DevLight L = a.base_lights[i];
    const double ss = a.theta_s[2 * a.n_groups + i], se = a.theta_e[2 * a.n_groups + i];
    for (int c = 0; c < 3; ++c) {
      a.rad_base[3 * i + c] = L.radiance[c];
      a.rad_e[3 * i + c] = se * L.radiance[c];
      L.radiance[c] = ss * L.radiance[c];
    }
    double x = fabs(L.bmax[0] - L.bmin[0]), y = fabs(L.bmax[1] - L.bmin[1]), z = fabs(L.bmax[2] - L.bmin[2]);
    double bbox_area = 2.0 * (x * y + x * z + y * z);
    double lum = 0.0;
    for (int c = 0; c < 3; ++c) lum += kLumW[c] * ((kPi * L.area) * L.radiance[c]);
    L.intensity = lum / (kPi * bbox_area);
    a.lights_s[i] = L;
  }
}

}  // namespace

hipError_t launch_dl_clear(const DlArgs &a, hipStream_t stream) {
  uint64_t n = (uint64_t)a.width * a.height * 3;
  uint32_t grid = (uint32_t)((n + 255) / 256);
  hipLaunchKernelGGL(k_dl_clear, dim3(grid > 0 ? grid : 1), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_dl_camera(const DlArgs &a, hipStream_t stream) {
  uint32_t grid = (uint32_t)((a.batch + 255) / 256);
  hipLaunchKernelGGL(k_dl_camera, dim3(grid > 0 ? grid : 1), dim3(256), 0, stream, a);
  return hipGetLastError();
}

template <bool SPLIT, bool ADJ>
static void launch_dl_samples(const DlArgs &a, hipStream_t stream) {
  const uint64_t nl = a.batch * a.nls, nb = a.batch * a.nbs, B = kSampleBlock;
  if (nl) hipLaunchKernelGGL((k_dl_light<SPLIT, ADJ>), dim3((uint32_t)((nl + B - 1) / B)), dim3(B), 0, stream, a);
  if (nb) hipLaunchKernelGGL((k_dl_bsdf<SPLIT, ADJ>), dim3((uint32_t)((nb + B - 1) / B)), dim3(B), 0, stream, a);
}

hipError_t launch_dl_point(const DlArgs &a, hipStream_t stream) {
  uint32_t grid = (uint32_t)((a.batch + 63) / 64);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_point, dim3(grid), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_dl_shade(const DlArgs &a, hipStream_t stream) {
  uint32_t grid = (uint32_t)((a.batch + 63) / 64);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_point, dim3(grid), dim3(64), 0, stream, a);
  if (a.adj || a.tape_j) launch_dl_samples<false, true>(a, stream);
  else if (a.mats_e != a.mats) launch_dl_samples<true, false>(a, stream);
  else launch_dl_samples<false, false>(a, stream);
  return hipGetLastError();
}

hipError_t launch_dl_resolve(const DlArgs &a, hipStream_t stream) {
  uint32_t grid = (uint32_t)((a.batch + 255) / 256);
  dim3 g(grid > 0 ? grid : 1);
  if (a.adj) hipLaunchKernelGGL((k_dl_resolve<1>), g, dim3(256), 0, stream, a);
  else if (a.tape_j) hipLaunchKernelGGL((k_dl_resolve<2>), g, dim3(256), 0, stream, a);
  else hipLaunchKernelGGL((k_dl_resolve<0>), g, dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_dl_tape_adjoint(const DlTapeArgs &t, hipStream_t stream) {
  uint64_t blocks = (t.n_samples + 255) / 256;
  uint32_t grid = (uint32_t)(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
  hipLaunchKernelGGL(k_dl_tape_adjoint, dim3(grid), dim3(256), 0, stream, t);
  return hipGetLastError();
}

hipError_t launch_dl_finish_f64(const DlArgs &a, double *img, hipStream_t stream) {
  uint32_t grid = (uint32_t)(((a.batch >> a.ppl_log2) + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_finish_f64, dim3(grid), dim3(256), 0, stream, a, img);
  return hipGetLastError();
}

hipError_t launch_dl_clear_grad(double *grad, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_dl_clear_grad, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1), dim3(256), 0, stream, grad, n);
  return hipGetLastError();
}

hipError_t launch_dl_params(const DlParamArgs &a, hipStream_t stream) {
  uint32_t n = a.n_prims > a.n_lights ? a.n_prims : a.n_lights;
  uint32_t grid = (n + 255) / 256;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_params, dim3(grid), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_pm_combine(const double *sample_dl, const double *sample_fg, uint64_t npix, uint32_t spp,
                             float *rgb, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((npix + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pm_combine, dim3(grid), dim3(256), 0, stream, sample_dl, sample_fg, npix, spp, rgb);
  return hipGetLastError();
}

hipError_t launch_dl_finish(const DlArgs &a, float *rgb, hipStream_t stream) {
  uint32_t grid = (uint32_t)(((a.batch >> a.ppl_log2) + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_finish, dim3(grid), dim3(256), 0, stream, a, rgb);
  return hipGetLastError();
}

hipError_t launch_dl_accum(const DlArgs &a, hipStream_t stream) {
  uint32_t grid = (uint32_t)(((a.batch >> a.ppl_log2) + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dl_accum, dim3(grid), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_fg_lookup_in(const FgArgs &f, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((f.a.batch + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fg_lookup_in, dim3(grid), dim3(256), 0, stream, f);
  return hipGetLastError();
}

hipError_t launch_fg_samples(const FgArgs &f, hipStream_t stream) {
  // slots_padded is a multiple of kFgBlock; at most 32 found photons per slot
  const uint64_t blocks = f.slots_padded / kFgBlock;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fg_samples, dim3((uint32_t)blocks), dim3(kFgSampleThreads), 0, stream, f);
  return hipGetLastError();
}

hipError_t launch_fg_quad(const FgArgs &f, hipStream_t stream) {
  hipLaunchKernelGGL(k_fg_quad, dim3(f.n_mats / 256 + 1), dim3(256), 0, stream, f);
  return hipGetLastError();
}

hipError_t launch_fg_hit(const FgArgs &f, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((f.slots_padded * 2ull * f.gs + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipError_t e = launch_set_counter(f.plastic_count, 0, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fg_hit, dim3(grid), dim3(256), 0, stream, f);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the list's length is read on the device: a grid of a few blocks per CU strides over it
  hipLaunchKernelGGL(k_fg_plastic, dim3(std::min<uint32_t>(grid, 4096)), dim3(256), 0, stream, f);
  return hipGetLastError();
}

hipError_t launch_fg_resolve(const FgArgs &f, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((f.a.batch + 255) / 256);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fg_resolve, dim3(grid), dim3(256), 0, stream, f);
  return hipGetLastError();
}

}  // namespace skw
