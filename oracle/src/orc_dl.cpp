// ORACLE — TEST INFRASTRUCTURE ONLY (see orc_math.hpp). Never linked into the product library.
//
// CPU restatement of Skwarka's direct-lighting render for scenes of matte / plastic surfaces lit
// by diffuse area lights, one camera sample at a time:
//   LTEIntegrator::Radiance                     Source/RayLibs/Raytracer/Core/LTEIntegrator.cpp:28-56
//   DirectLightingLTEIntegrator::_SurfaceRadiance  Raytracer/LTEIntegrators/DirectLightingLTEIntegrator.cpp:33-60
//   DirectLightingIntegrator::ComputeDirectLighting / _SampleLights / _SampleBSDF
//                                               Raytracer/Core/DirectLightingIntegrator.cpp:54-366
//   BSDF::Evaluate / PDF / Sample, frame        Raytracer/Core/BSDF.cpp:22-200, BSDF.h:182-244
//   BxDF::Sample / PDF (cosine hemisphere)      Raytracer/Core/BxDF.cpp:20-41; Lambertian.cpp:18-31
//   Microfacet<FresnelDielectric, Blinn>        Raytracer/BxDFs/Microfacet.h:89-152
//   BlinnDistribution                           Raytracer/MicrofacetDistributions/BlinnDistribution.h:70-110
//   FresnelDielectric                           Raytracer/Core/Fresnel.h:92-124
//   MatteMaterial (sigma = 0) / PlasticMaterial Raytracer/Materials/MatteMaterial.cpp:29-48, PlasticMaterial.cpp:33-52
//   DiffuseAreaLightSource                      Raytracer/LightSources/DiffuseAreaLightSource.cpp:20-174
//   IrradianceLightsSamplingStrategy            Raytracer/LightsSamplingStrategies/IrradianceLightsSamplingStrategy.cpp:19-129
//   SubtendedSolidAngle, CoordinateSystem, BinarySearchCDF  Source/RayLibs/Math/MathRoutines.h:192-215, 298-390
//   ConcentricDiskSampling, CosineHemisphereSampling, UniformTriangleSampling, PowerHeuristic
//                                               Source/RayLibs/Math/SamplingRoutines.h:197-325, 449-455
//   TriangleMesh::GetTriangleNormal, Triangle3D::GetArea   TriangleMesh.h:219-232, Math/Triangle3D.h:82-86
//   PerspectiveCamera::GenerateRay              Raytracer/Cameras/PerspectiveCamera.cpp:36-64
//   NaN / negative sample clamp                 Raytracer/Renderers/SamplerBasedRenderer.cpp:459-478
// Spectrum luminance weights are the Y row of the reference's spectrum-to-XYZ matrix
// (SpectrumRoutines.cpp:20-41 summed over its CIE Y table for the bands 360-494, 495-574, 575-830 nm).
// Samples come from a counter-based stream (splitmix64 of seed, sample index and dimension) in
// place of the reference's samplers; the product reproduces the same stream.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "../include/orc_api.h"
#include "orc_math.hpp"

namespace orc {

// Spectrum_d luminance: Y row of m_Spectrum_to_XYZ (SpectrumRoutines.cpp:55-62).
static const double LUM_W[3] = {0x1.b38293da589a6p-5, 0x1.2b993cc221566p-1, 0x1.725d3400723f6p-2};
static inline double luminance(const double s[3]) {
  double r = 0.0;
  for (int j = 0; j < 3; ++j) r += LUM_W[j] * s[j];
  return r;
}

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Uniform double in [0, 1) for (seed, sample index, dimension).
static inline double rnd(uint64_t seed, uint64_t sample, uint32_t dim) {
  uint64_t h = mix64(seed ^ mix64(sample * 64ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}

// Camera and material description: see orc_api.h.
struct Light {
  uint32_t prim;
  double radiance[3];
  std::vector<double> cdf;       // normalised area CDF over the light mesh's triangles
  std::vector<V3> normal;        // float GetTriangleNormal, as double
  std::vector<V3> v0, v1, v2;    // float vertices as double
  double area = 0.0;
  BBoxD bbox;
  double intensity = 0.0;
};

struct DLScene {
  std::vector<uint8_t> mat;
  std::vector<double> kd, ks, rough;
  std::vector<Light> lights;
  std::vector<int> light_of_prim;
  uint32_t n_groups = 0;
  std::vector<uint32_t> group;  // per prim, NO_GROUP if none
};
static const uint32_t NO_GROUP = 0xFFFFFFFFu;

// Triangle3D<float>::GetArea: float cross product, float length, (float)0.5 * length.
static float tri_area_f(F3 a, F3 b, F3 c) {
  float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
  float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
  float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
  return 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
}
// TriangleMesh::GetTriangleNormal: float cross product normalised in float (Vector3D<float>).
static V3 tri_normal_f(F3 a, F3 b, F3 c, bool invert) {
  float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
  float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
  float cx, cy, cz;
  if (invert) {
    cx = e2y * e1z - e2z * e1y;
    cy = e2z * e1x - e2x * e1z;
    cz = e2x * e1y - e2y * e1x;
  } else {
    cx = e1y * e2z - e1z * e2y;
    cy = e1z * e2x - e1x * e2z;
    cz = e1x * e2y - e1y * e2x;
  }
  float l2 = cx * cx + cy * cy + cz * cz;
  if (l2 > 0.0f) {
    double inv = 1.0 / (double)std::sqrt(l2);
    cx = (float)(cx * inv);
    cy = (float)(cy * inv);
    cz = (float)(cz * inv);
  }
  return V3{cx, cy, cz};
}

// MathRoutines::_SubtendedSolidAngle / SubtendedSolidAngle
static double solid_angle3(V3 d1, V3 d2, V3 d3) {
  double det = dot(d1, cross(d2, d3));
  double denom = 1 + dot(d1, d2) + dot(d1, d3) + dot(d2, d3);
  if (denom > 0.0) return std::fabs(std::atan(det / denom) * 2.0);
  if (denom < 0.0) return 2.0 * (PI_ - std::fabs(std::atan(det / denom)));
  return 2.0 * PI_;
}
static double subtended_solid_angle(V3 p, const BBoxD &b) {
  V3 dirs[8];
  for (int i = 0; i < 8; ++i) {
    V3 v{(i & 1) == 0 ? b.mn[0] : b.mx[0], (i & 2) == 0 ? b.mn[1] : b.mx[1], (i & 4) == 0 ? b.mn[2] : b.mx[2]};
    V3 d = sub(v, p);
    normalize(d);
    dirs[i] = d;
  }
  bool inside = true;
  double sum = 0.0;
  if (p.x < b.mn[0] - DBL_EPS_) {
    sum += solid_angle3(dirs[0], dirs[2], dirs[6]) + solid_angle3(dirs[0], dirs[4], dirs[6]);
    inside = false;
  } else if (p.x > b.mx[0] + DBL_EPS_) {
    sum += solid_angle3(dirs[1], dirs[3], dirs[7]) + solid_angle3(dirs[1], dirs[5], dirs[7]);
    inside = false;
  }
  if (p.y < b.mn[1] - DBL_EPS_) {
    sum += solid_angle3(dirs[0], dirs[1], dirs[5]) + solid_angle3(dirs[0], dirs[4], dirs[5]);
    inside = false;
  } else if (p.y > b.mx[1] + DBL_EPS_) {
    sum += solid_angle3(dirs[2], dirs[3], dirs[7]) + solid_angle3(dirs[2], dirs[6], dirs[7]);
    inside = false;
  }
  if (p.z < b.mn[2] - DBL_EPS_) {
    sum += solid_angle3(dirs[0], dirs[1], dirs[3]) + solid_angle3(dirs[0], dirs[2], dirs[3]);
    inside = false;
  } else if (p.z > b.mx[2] + DBL_EPS_) {
    sum += solid_angle3(dirs[4], dirs[5], dirs[7]) + solid_angle3(dirs[4], dirs[6], dirs[7]);
    inside = false;
  }
  return inside ? 4.0 * PI_ : sum;
}

static bool bbox_inside(const BBoxD &b, V3 p) {
  return p.x >= b.mn[0] && p.y >= b.mn[1] && p.z >= b.mn[2] && p.x <= b.mx[0] && p.y <= b.mx[1] && p.z <= b.mx[2];
}

// IrradianceLightsSamplingStrategy::GetLightsCDF(point, normal) for area lights only.
static void lights_cdf(const DLScene &s, const double *intensity, V3 p, V3 n, double *cdf) {
  size_t nl = s.lights.size();
  for (size_t i = 0; i < nl; ++i) {
    const Light &L = s.lights[i];
    if (bbox_inside(L.bbox, p)) {
      cdf[i] = PI_ * intensity[i];
    } else {
      double avg = 0.0;
      for (int j = 0; j < 8; ++j) {
        V3 v{(j & 1) ? L.bbox.mn[0] : L.bbox.mx[0], (j & 2) ? L.bbox.mn[1] : L.bbox.mx[1],
             (j & 4) ? L.bbox.mn[2] : L.bbox.mx[2]};
        V3 d = sub(v, p);
        normalize(d);
        avg += smax(0.0, dot(d, n));
      }
      avg *= 0.125;
      if (avg > 0.0) {
        if (avg < 0.1) avg = 0.1;
        cdf[i] = subtended_solid_angle(p, L.bbox) * intensity[i] * avg;
      } else {
        cdf[i] = 0.0;
      }
    }
  }
  for (size_t i = 1; i < nl; ++i) cdf[i] += cdf[i - 1];
  if (cdf[nl - 1] > 0.0) {
    double inv = 1.0 / cdf[nl - 1];
    for (size_t i = 0; i < nl; ++i) cdf[i] *= inv;
  } else {
    double inv = 1.0 / (double)nl;
    for (size_t i = 0; i < nl; ++i) cdf[i] = ((double)i + 1.0) * inv;
  }
}

// DiffuseAreaLightSource::SampleLighting (DiffuseAreaLightSource.cpp:73-102): triangle by the area
// CDF, point by UniformTriangleSampling (SamplingRoutines.h:318-325), pdf with respect to solid
// angle = d^2 / (area |cos|); radiance only toward the lit side (normal . wi < 0).
static size_t search_cdf(const double *b, size_t n, double u, double *pdf);
static void light_sample(const Light &Lt, V3 p, double tri_u, double u0, double u1, size_t &ti, V3 &wi, double &len,
                         double &lpdf, bool &lit) {
  ti = search_cdf(Lt.cdf.data(), Lt.cdf.size(), tri_u, nullptr);
  double tmp = std::sqrt(u0);
  double bb1 = 1.0 - tmp, bb2 = u1 * tmp;
  V3 sp = add(add(mul(Lt.v0[ti], bb1), mul(Lt.v1[ti], bb2)), mul(Lt.v2[ti], 1.0 - bb1 - bb2));
  V3 lv = sub(sp, p);
  len = std::sqrt(len_sqr(lv));
  wi = mul(lv, 1.0 / len);
  double ac = std::fabs(dot(wi, Lt.normal[ti]));
  lpdf = ac > DBL_EPS_ ? (len * len) / (Lt.area * ac) : 0.0;
  lit = dot(Lt.normal[ti], wi) < 0.0;
}
// DiffuseAreaLightSource::LightingPDF (DiffuseAreaLightSource.cpp:104-116) of a ray that reaches
// triangle tri after distance dist.
static double light_pdf(const Light &Lt, size_t tri, V3 wi, double dist) {
  double ac = std::fabs(dot(wi, Lt.normal[tri]));
  return ac > DBL_EPS_ ? (dist * dist) / (Lt.area * ac) : 0.0;
}

// MathRoutines::BinarySearchCDF
static size_t search_cdf(const double *b, size_t n, double u, double *pdf) {
  size_t lo = 0, hi = n - 1;
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (b[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  if (pdf) *pdf = lo == 0 ? b[0] : b[lo] - b[lo - 1];
  return lo;
}

// SamplingRoutines::ConcentricDiskSampling
static void concentric_disk(double u0, double u1, double &x, double &y) {
  double s0 = 2.0 * u0 - 1.0, s1 = 2.0 * u1 - 1.0;
  if (s0 == 0.0 && s1 == 0.0) {
    x = y = 0.0;
    return;
  }
  double r, theta;
  if (s0 >= -s1) {
    if (s0 > s1) {
      r = s0;
      theta = s1 > 0.0 ? s1 / r : 8.0 + s1 / r;
    } else {
      r = s1;
      theta = 2.0 - s0 / r;
    }
  } else {
    if (s0 <= s1) {
      r = -s0;
      theta = 4.0 - s1 / r;
    } else {
      r = -s1;
      theta = 6.0 + s0 / r;
    }
  }
  theta *= PI_ / 4.0;
  x = r * std::cos(theta);
  y = r * std::sin(theta);
}

// FresnelDielectric(1.5, 1.0)
static double fresnel_dielectric(double c) {
  double ni = 1.5, no = 1.0;
  c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
  if (c < 0.0) std::swap(ni, no);
  double sin_t = no / ni * std::sqrt(smax(0.0, 1.0 - c * c));
  if (sin_t >= 1.0) return 1.0;
  double cos_t = std::sqrt(smax(0.0, 1.0 - sin_t * sin_t));
  c = std::fabs(c);
  double rpar = ((ni * c) - (no * cos_t)) / ((ni * c) + (no * cos_t));
  double rper = ((no * c) - (ni * cos_t)) / ((no * c) + (ni * cos_t));
  return (rpar * rpar + rper * rper) * 0.5;
}

struct Bsdf {
  V3 n, ng, e1, e2;
  bool plastic;
  double kd[3], ks[3];
  double e;  // Blinn exponent

  V3 to_local(V3 v) const { return V3{dot(v, e1), dot(v, e2), dot(v, n)}; }
  V3 to_world(V3 v) const { return add(add(mul(e1, v.x), mul(e2, v.y)), mul(n, v.z)); }

  double blinn_d(V3 h) const { return (e + 2.0) * 0.15915494309189533577 * std::pow(std::fabs(h.z), e); }
  double blinn_pdf(V3 wo, V3 wi) const {
    V3 h = add(wo, wi);
    normalize(h);
    double ct = std::fabs(h.z);
    return ((e + 1.0) * std::pow(ct, e)) / (2.0 * PI_ * 4.0 * std::fabs(dot(wo, h)));
  }
  // Microfacet::Evaluate(incident, exitant), scalar part (multiplied by ks per channel)
  bool micro_eval(V3 wo, V3 wi, double &scalar, double &fres) const {
    double ci = std::fabs(wo.z), co = std::fabs(wi.z);
    if (ci < DBL_EPS_ || co < DBL_EPS_) return false;
    V3 h = add(wo, wi);
    if (std::fabs(h.x) < DBL_EPS_ && std::fabs(h.y) < DBL_EPS_ && std::fabs(h.z) < DBL_EPS_) return false;
    normalize(h);
    fres = fresnel_dielectric(dot(wi, h));
    double ch = std::fabs(h.z);
    double tmp = 2.0 * ch / std::fabs(dot(wo, h));
    double g = smin(1.0, smin(ci * tmp, co * tmp));
    scalar = blinn_d(h) * g / (4.0 * ci * co);
    return true;
  }
  double micro_pdf(V3 wo, V3 wi) const {
    if (wo.z * wi.z < 0.0) return 0.0;
    return blinn_pdf(wo, wi);
  }
  static double lambert_pdf(V3 wo, V3 wi) {
    bool same = wo.z * wi.z > 0.0;
    return same ? std::fabs(wi.z) * INV_PI_ : 0.0;
  }
  // Sum of the reflection BxDFs (all components here are reflective), local vectors.
  void eval_reflect_local(V3 wo, V3 wi, double f[3]) const {
    for (int c = 0; c < 3; ++c) f[c] = kd[c] * INV_PI_;
    if (plastic) {
      double sc, fr;
      if (micro_eval(wo, wi, sc, fr))
        for (int c = 0; c < 3; ++c) f[c] += (ks[c] * fr) * sc;
    }
  }
  // BSDF::Evaluate(incident, exitant)
  void evaluate(V3 wo_w, V3 wi_w, double f[3]) const {
    V3 wo = to_local(wo_w), wi = to_local(wi_w);
    if (dot(wo_w, ng) * dot(wi_w, ng) > 0) {
      eval_reflect_local(wo, wi, f);
    } else {
      f[0] = f[1] = f[2] = 0.0;
    }
  }
  // BSDF::PDF(incident, exitant) over all components
  double pdf(V3 wo_w, V3 wi_w) const {
    V3 wo = to_local(wo_w), wi = to_local(wi_w);
    double p = lambert_pdf(wo, wi);
    if (!plastic) return p;
    p += micro_pdf(wo, wi);
    return p / 2.0;
  }
  // BSDF::Sample with non-specular flags.
  // wo_l / wi_l: the local pair f was evaluated with.
  bool sample(V3 wo_w, double u0, double u1, double uc, V3 &wi_w, double &pdf_out, double f[3], V3 &wo_l,
              V3 &wi_l) const {
    size_t num = plastic ? 2 : 1;
    size_t comp = smin((size_t)(uc * (double)num), num - 1);
    V3 wo = to_local(wo_w);
    V3 wi;
    double p = 0.0;
    if (comp == 0) {
      double dx, dy;
      concentric_disk(u0, u1, dx, dy);
      double z = std::sqrt(smax(0.0, 1.0 - dx * dx - dy * dy));
      wi = V3{dx, dy, z};
      if (wo.z < 0.0) wi.z *= -1.0;
      p = lambert_pdf(wo, wi);
    } else {
      double ct = std::pow(u0, 1.0 / (e + 1.0));
      double st = std::sqrt(smax(0.0, 1.0 - ct * ct));
      double phi = u1 * 2.0 * PI_;
      V3 h{st * std::cos(phi), st * std::sin(phi), ct};
      if (wo.z * h.z < 0.0) h.z *= -1.0;
      wi = sub(mul(h, 2.0 * dot(wo, h)), wo);
      p = ((e + 1.0) * std::pow(ct, e)) / (2.0 * PI_ * 4.0 * std::fabs(dot(wo, h)));
      if (wo.z * wi.z < 0.0) p = 0.0;
    }
    wi_w = to_world(wi);
    if (num > 1) {
      p += comp == 0 ? micro_pdf(wo, wi) : lambert_pdf(wo, wi);
      p /= (double)num;
    }
    pdf_out = p;
    wo_l = wo;
    wi_l = wi;
    if (dot(wo_w, ng) * dot(wi_w, ng) > 0) eval_reflect_local(wo, wi, f);
    else f[0] = f[1] = f[2] = 0.0;
    return true;
  }
};

static inline double power_heuristic(double n1, double p1, double n2, double p2) {
  double f = n1 * p1, g = n2 * p2;
  return (f * f) / (f * f + g * g);
}

// ---------------------------------------------------------------- forward-mode derivatives
// Dual number over the parameters one camera sample can depend on: [0] reflectance of the hit
// primitive's group, [1] its roughness, [2 + l] the emission scale of light l. Values follow
// exactly the double operations of the plain render, so value parts are bit-identical to it.
constexpr int ND = 2 + 16;
struct Dual {
  double v = 0.0;
  double d[ND] = {};
  Dual() = default;
  Dual(double x) : v(x) {}
};
static inline Dual operator+(const Dual &a, const Dual &b) {
  Dual r(a.v + b.v);
  for (int j = 0; j < ND; ++j) r.d[j] = a.d[j] + b.d[j];
  return r;
}
static inline Dual operator+(const Dual &a, double b) {
  Dual r = a;
  r.v = a.v + b;
  return r;
}
static inline Dual operator*(const Dual &a, const Dual &b) {
  Dual r(a.v * b.v);
  for (int j = 0; j < ND; ++j) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
static inline Dual operator*(const Dual &a, double b) {
  Dual r(a.v * b);
  for (int j = 0; j < ND; ++j) r.d[j] = a.d[j] * b;
  return r;
}
static inline Dual operator*(double a, const Dual &b) {
  Dual r(a * b.v);
  for (int j = 0; j < ND; ++j) r.d[j] = a * b.d[j];
  return r;
}
static inline Dual operator/(const Dual &a, double b) {
  Dual r(a.v / b);
  for (int j = 0; j < ND; ++j) r.d[j] = a.d[j] / b;
  return r;
}
static inline double value(double x) { return x; }
static inline double value(const Dual &x) { return x.v; }
// b^e for a constant base b >= 0 (d/de b^e = b^e ln b, taken as 0 where b^e is 0).
static inline double pow_c(double b, double e) { return std::pow(b, e); }
static inline Dual pow_c(double b, const Dual &e) {
  Dual r(std::pow(b, e.v));
  double k = r.v > 0.0 ? r.v * std::log(b) : 0.0;
  for (int j = 0; j < ND; ++j) r.d[j] = k * e.d[j];
  return r;
}
template <class T>
static inline T seeded(double v, int dir, double dv);
template <>
inline double seeded<double>(double v, int, double) {
  return v;
}
template <>
inline Dual seeded<Dual>(double v, int dir, double dv) {
  Dual r(v);
  r.d[dir] = dv;
  return r;
}

// Evaluation-side material: the values the estimator is differentiated through.
template <class T>
struct EvalMat {
  T kd[3];
  double ks[3];
  T e;
  bool plastic;
};

// Bsdf::eval_reflect_local with the evaluation material.
template <class T>
static void eval_local(const EvalMat<T> &m, V3 wo, V3 wi, T f[3]) {
  for (int c = 0; c < 3; ++c) f[c] = m.kd[c] * INV_PI_;
  if (!m.plastic) return;
  double ci = std::fabs(wo.z), co = std::fabs(wi.z);
  if (ci < DBL_EPS_ || co < DBL_EPS_) return;
  V3 h = add(wo, wi);
  if (std::fabs(h.x) < DBL_EPS_ && std::fabs(h.y) < DBL_EPS_ && std::fabs(h.z) < DBL_EPS_) return;
  normalize(h);
  double fres = fresnel_dielectric(dot(wi, h));
  double ch = std::fabs(h.z);
  double tmp = 2.0 * ch / std::fabs(dot(wo, h));
  double g = smin(1.0, smin(ci * tmp, co * tmp));
  T d = (m.e + 2.0) * 0.15915494309189533577 * pow_c(ch, m.e);
  T sc = d * g / (4.0 * ci * co);
  for (int c = 0; c < 3; ++c) f[c] = f[c] + (m.ks[c] * fres) * sc;
}
// Bsdf::evaluate (world vectors) with the evaluation material.
template <class T>
static void eval_world(const Bsdf &frame, const EvalMat<T> &m, V3 wo_w, V3 wi_w, T f[3]) {
  V3 wo = frame.to_local(wo_w), wi = frame.to_local(wi_w);
  if (dot(wo_w, frame.ng) * dot(wi_w, frame.ng) > 0) {
    eval_local(m, wo, wi, f);
  } else {
    for (int c = 0; c < 3; ++c) f[c] = 0.0;
  }
}

// Material and light values of one parameter vector. A grouped primitive takes gray kd =
// clamp(theta[g], 0, 1) and exponent 1/clamp(theta[G+g], 0.001, 1) clamped to [0, 1000]; light l
// takes radiance theta[2G+l] * its base radiance. Derivatives are 0 where a clamp holds.
struct ParamView {
  std::vector<double> kd, e, dkd, de_dr;  // kd: 3 per prim
  std::vector<double> rad, intensity;     // rad: 3 per light
};

static void make_view(const DLScene &s, const double *theta, ParamView &v) {
  const size_t np = s.mat.size(), nl = s.lights.size();
  const uint32_t G = s.n_groups;
  v.kd.assign(3 * np, 0.0);
  v.e.assign(np, 0.0);
  v.dkd.assign(np, 0.0);
  v.de_dr.assign(np, 0.0);
  for (size_t p = 0; p < np; ++p) {
    for (int c = 0; c < 3; ++c) v.kd[3 * p + c] = smin(1.0, smax(0.0, s.kd[3 * p + c]));
    double rough = s.rough[p] < 0.001 ? 0.001 : (s.rough[p] > 1.0 ? 1.0 : s.rough[p]);
    double ex = 1.0 / rough;
    v.e[p] = ex < 0.0 ? 0.0 : (ex > 1000.0 ? 1000.0 : ex);
    if (theta && s.group[p] != NO_GROUP) {
      const uint32_t g = s.group[p];
      const double rho = theta[g];
      const double kd = rho < 0.0 ? 0.0 : (rho > 1.0 ? 1.0 : rho);
      for (int c = 0; c < 3; ++c) v.kd[3 * p + c] = kd;
      v.dkd[p] = (rho > 0.0 && rho < 1.0) ? 1.0 : 0.0;
      const double r = theta[G + g];
      const double rc = r < 0.001 ? 0.001 : (r > 1.0 ? 1.0 : r);
      const double e = 1.0 / rc;
      v.e[p] = e < 0.0 ? 0.0 : (e > 1000.0 ? 1000.0 : e);
      v.de_dr[p] = (r > 0.001 && r < 1.0) ? -1.0 / (r * r) : 0.0;
    }
  }
  v.rad.assign(3 * nl, 0.0);
  v.intensity.assign(nl, 0.0);
  for (size_t l = 0; l < nl; ++l) {
    const Light &L = s.lights[l];
    if (!theta) {
      for (int c = 0; c < 3; ++c) v.rad[3 * l + c] = L.radiance[c];
      v.intensity[l] = L.intensity;
      continue;
    }
    const double sc = theta[2 * G + l];
    double power[3];
    for (int c = 0; c < 3; ++c) {
      v.rad[3 * l + c] = sc * L.radiance[c];
      power[c] = (PI_ * L.area) * v.rad[3 * l + c];
    }
    v.intensity[l] = luminance(power) / (PI_ * L.bbox.area());
  }
}

struct FrameSpec {
  const orc_accel *accel;
  const DLScene *s;
  const double *cam;
  uint32_t width, height, spp, nls, nbs;
  uint64_t seed;
  const ParamView *vs;  // sampling: BSDF sampling, pdfs, light CDF, control flow
  const ParamView *ve;  // evaluation: BSDF values and emitted radiance
};

// One camera sample of LTEIntegrator::Radiance with direct lighting (the per-sample body of the
// render), with values of type T taken from the evaluation view. Returns whether the sample
// counts (the SamplerBasedRenderer rule) and the primitive the camera ray hit.
template <class T>
static bool dl_sample(const FrameSpec &F, uint64_t pix, uint32_t si, T L[3], uint32_t &hit_prim, uint32_t &hit_tri,
                      uint64_t &c_shadow, uint64_t &c_bsdf) {
  const DLScene &s = *F.s;
  const ParamView &vs = *F.vs, &ve = *F.ve;
  const double *cam = F.cam;
  const size_t nl = s.lights.size();
  const uint32_t px = (uint32_t)(pix % F.width), py = (uint32_t)(pix / F.width);
  const uint64_t sidx = pix * F.spp + si;
  // PerspectiveCamera::GenerateRay with lens radius 0
  double ix = (double)px + rnd(F.seed, sidx, 0), iy = (double)py + rnd(F.seed, sidx, 1);
  double cx = 2.0 * cam[12] * (ix - F.width * 0.5) / F.width;
  double cy = 2.0 * cam[13] * (iy - F.height * 0.5) / F.height;
  V3 d{cam[3] * cx + cam[4] * cy + cam[5], cam[6] * cx + cam[7] * cy + cam[8], cam[9] * cx + cam[10] * cy + cam[11]};
  normalize(d);
  double ray[8] = {cam[0], cam[1], cam[2], d.x, d.y, d.z, 0.0, DBL_INF_};
  orc_hit h;
  orc_hit_geom g;
  orc_accel_intersect_one(F.accel, ray, &h, &g);
  hit_prim = h.prim;
  hit_tri = h.tri;
  for (int c = 0; c < 3; ++c) L[c] = 0.0;
  if (h.prim != 0xffffffffu) {
    V3 wo = mul(d, -1.0);
    V3 p{g.p[0], g.p[1], g.p[2]}, ng{g.ng[0], g.ng[1], g.ng[2]}, ns{g.ns[0], g.ns[1], g.ns[2]};
    T Le[3] = {0.0, 0.0, 0.0};
    int li = s.light_of_prim[h.prim];
    if (li >= 0 && dot(wo, ng) > 0.0)
      for (int c = 0; c < 3; ++c) Le[c] = seeded<T>(ve.rad[3 * li + c], 2 + li, s.lights[li].radiance[c]);
    // BSDF frame (BSDF.h:182-192) with the DG tangent of a mesh without UVs or tangents
    // (TriangleMesh.cpp:520-540): CoordinateSystem(ns), then made orthogonal to ns.
    V3 t;
    if (std::fabs(ns.x) > std::fabs(ns.y)) {
      double il = 1.0 / std::sqrt(ns.x * ns.x + ns.z * ns.z);
      t = V3{-ns.z * il, 0.0, ns.x * il};
    } else {
      double il = 1.0 / std::sqrt(ns.y * ns.y + ns.z * ns.z);
      t = V3{0.0, ns.z * il, -ns.y * il};
    }
    V3 t2 = cross(ns, t);
    V3 tangent;
    if (len_sqr(t2) > DBL_EPS_) {
      tangent = cross(t2, ns);
      normalize(tangent);
    } else {
      tangent = t;
    }
    const uint32_t pr = h.prim;
    Bsdf b;  // sampling BSDF
    b.n = ns;
    b.ng = ng;
    b.e1 = tangent;
    b.e2 = cross(ns, tangent);
    b.plastic = s.mat[pr] == 1;
    for (int c = 0; c < 3; ++c) {
      b.kd[c] = vs.kd[3 * pr + c];
      b.ks[c] = smin(1.0, smax(0.0, s.ks[3 * pr + c]));
    }
    b.e = vs.e[pr];
    EvalMat<T> me;  // evaluation material
    me.plastic = b.plastic;
    for (int c = 0; c < 3; ++c) {
      me.kd[c] = seeded<T>(ve.kd[3 * pr + c], 0, ve.dkd[pr]);
      me.ks[c] = b.ks[c];
    }
    me.e = seeded<T>(ve.e[pr], 1, ve.de_dr[pr]);
    auto rad_e = [&](size_t l, int c) { return seeded<T>(ve.rad[3 * l + c], 2 + (int)l, s.lights[l].radiance[c]); };
    T cdl[3] = {0.0, 0.0, 0.0};
    if (nl > 0) {
      std::vector<double> cdf(nl);
      V3 nrm = dot(wo, ns) >= 0.0 ? ns : mul(ns, -1.0);
      lights_cdf(s, vs.intensity.data(), p, nrm, cdf.data());
      T dl[3] = {0.0, 0.0, 0.0};
      // _SampleLights (area lights only)
      for (uint32_t k = 0; k < F.nls; ++k) {
        double uc = rnd(F.seed, sidx, 2 + k);
        double u0 = rnd(F.seed, sidx, 6 + 2 * k), u1 = rnd(F.seed, sidx, 7 + 2 * k);
        double comp_pdf;
        size_t l = search_cdf(cdf.data(), nl, uc, &comp_pdf);
        const Light &Lt = s.lights[l];
        double tri_u = (l == 0 ? uc : uc - cdf[l - 1]) / comp_pdf;
        if (Lt.cdf.empty()) continue;
        size_t ti;
        V3 wi;
        double len, lpdf;
        bool lit;
        light_sample(Lt, p, tri_u, u0, u1, ti, wi, len, lpdf, lit);
        const double *rs = &vs.rad[3 * l];
        double le[3] = {lit ? rs[0] : 0.0, lit ? rs[1] : 0.0, lit ? rs[2] : 0.0};
        if (!(lpdf > 0.0) || (le[0] == 0.0 && le[1] == 0.0 && le[2] == 0.0)) continue;
        double f[3];
        b.evaluate(wo, wi, f);
        double lf[3] = {le[0] * f[0], le[1] * f[1], le[2] * f[2]};
        if (lf[0] == 0.0 && lf[1] == 0.0 && lf[2] == 0.0) continue;
        double mint = orc_next_min_t(F.accel, pr, &g, &wi.x);
        double sray[8] = {p.x, p.y, p.z, wi.x, wi.y, wi.z, mint, len - 1e-4};
        uint8_t occ = 0;
        orc_any_hit(F.accel, 1, sray, &occ, 1);
        ++c_shadow;
        if (occ) continue;
        double bp = b.pdf(wo, wi);
        double w = power_heuristic((double)F.nls, lpdf * comp_pdf, (double)F.nbs, bp);
        w *= std::fabs(dot(wi, ns)) / (lpdf * comp_pdf);
        T fe[3];
        eval_world(b, me, wo, wi, fe);
        for (int c = 0; c < 3; ++c) {
          T v = (lit ? rad_e(l, c) : T(0.0)) * fe[c];
          dl[c] = dl[c] + v * w;
        }
      }
      // ComputeDirectLighting = (0 + _SampleLights) + _SampleBSDF; each divides by its sample
      // count through Spectrum::operator/ (multiplication by the reciprocal).
      for (int c = 0; c < 3; ++c) cdl[c] = dl[c] * (1.0 / (double)F.nls);
      // _SampleBSDF
      T db[3] = {0.0, 0.0, 0.0};
      for (uint32_t k = 0; k < F.nbs; ++k) {
        double uc = rnd(F.seed, sidx, 14 + k);
        double u0 = rnd(F.seed, sidx, 18 + 2 * k), u1 = rnd(F.seed, sidx, 19 + 2 * k);
        V3 wi, wo_l, wi_l;
        double bp, f[3];
        b.sample(wo, u0, u1, uc, wi, bp, f, wo_l, wi_l);
        if (!(bp > 0.0) || (f[0] == 0.0 && f[1] == 0.0 && f[2] == 0.0)) continue;
        double mint = orc_next_min_t(F.accel, pr, &g, &wi.x);
        double bray[8] = {p.x, p.y, p.z, wi.x, wi.y, wi.z, mint, DBL_INF_};
        orc_hit h2;
        orc_hit_geom g2;
        orc_accel_intersect_one(F.accel, bray, &h2, &g2);
        ++c_bsdf;
        if (h2.prim == 0xffffffffu) continue;
        int l2 = s.light_of_prim[h2.prim];
        if (l2 < 0) continue;
        V3 ng2{g2.ng[0], g2.ng[1], g2.ng[2]};
        if (!(dot(mul(wi, -1.0), ng2) > 0.0)) continue;
        const Light &Lt = s.lights[l2];
        const double *rs = &vs.rad[3 * l2];
        if (rs[0] == 0.0 && rs[1] == 0.0 && rs[2] == 0.0) continue;
        double lpdf = light_pdf(Lt, h2.tri, wi, h2.t - mint);
        double cp = l2 == 0 ? cdf[0] : cdf[l2] - cdf[l2 - 1];
        double w = power_heuristic((double)F.nbs, bp, (double)F.nls, lpdf * cp);
        w *= std::fabs(dot(wi, ns)) / bp;
        // f is non-black, so the pair passed the geometric-normal test: evaluate it again with the
        // evaluation material.
        T fe[3];
        eval_local(me, wo_l, wi_l, fe);
        for (int c = 0; c < 3; ++c) db[c] = db[c] + fe[c] * rad_e(l2, c) * w;
      }
      for (int c = 0; c < 3; ++c) cdl[c] = cdl[c] + db[c] * (1.0 / (double)F.nbs);
    }
    for (int c = 0; c < 3; ++c) L[c] = Le[c] + cdl[c];
  }
  // SamplerBasedRenderer.cpp:458-478: NaN, negative-luminance and infinite samples become black
  double Lv[3] = {value(L[0]), value(L[1]), value(L[2])};
  bool bad = std::isnan(Lv[0]) || std::isnan(Lv[1]) || std::isnan(Lv[2]) || luminance(Lv) < -DBL_EPS_ ||
             std::isinf(Lv[0]) || std::isinf(Lv[1]) || std::isinf(Lv[2]);
  return !bad;
}

// ------------------------------------------------------------------ final gather

// Counter-based stream of the final gather's stratified samples (a stream of its own: dims up to
// 4096 per camera sample).
static inline double fg_rnd(uint64_t seed, uint64_t sample, uint32_t dim) {
  uint64_t h = mix64((seed ^ 0x6A09E667F3BCC909ull) ^ mix64(sample * 4096ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}

static const double FG_COS_GATHER = 0.9848;  // PhotonLTEIntegrator.cpp:396 (10 degrees)

// MathRoutines::CoordinateSystem (MathRoutines.h:298-313)
static void coordinate_system(V3 e1, V3 &e2, V3 &e3) {
  if (std::fabs(e1.x) > std::fabs(e1.y)) {
    double inv = 1.0 / std::sqrt(e1.x * e1.x + e1.z * e1.z);
    e2 = V3{-e1.z * inv, 0.0, e1.x * inv};
  } else {
    double inv = 1.0 / std::sqrt(e1.y * e1.y + e1.z * e1.z);
    e2 = V3{0.0, e1.z * inv, -e1.y * inv};
  }
  e3 = cross(e1, e2);
}

// BSDF::TotalScattering(incident, 5x5 stratified samples, BSDF_ALL_REFLECTION) of a matte /
// plastic surface (BSDF.cpp:209-224): Lambertian::TotalScattering = R (Lambertian.cpp:32-36) plus,
// for plastic, BxDF::TotalScattering of the microfacet lobe (BxDF.cpp:43-68) with
// Microfacet::Sample (Microfacet.h:112-128) / BlinnDistribution::Sample (BlinnDistribution.h:81-98).
// BSDF::TotalScattering hands the BxDFs the world-space incident vector (BSDF.cpp:213-219), which is
// kept here: the lobe is integrated around the world z axis.
static void total_scattering(const Bsdf &b, V3 incident, double ts[3]) {
  double r[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 3; ++c) r[c] = r[c] + b.kd[c];
  if (b.plastic) {
    double m[3] = {0.0, 0.0, 0.0};
    for (int y = 0; y < 5; ++y)
      for (int x = 0; x < 5; ++x) {
        const double u0 = (x + 0.5) * (1.0 / 5), u1 = (y + 0.5) * (1.0 / 5);
        const double ct = std::pow(u0, 1.0 / (b.e + 1.0));
        const double st = std::sqrt(smax(0.0, 1.0 - ct * ct));
        const double phi = u1 * 2.0 * PI_;
        V3 h{st * std::cos(phi), st * std::sin(phi), ct};
        if (incident.z * h.z < 0.0) h.z *= -1.0;
        V3 ex = sub(mul(h, 2.0 * dot(incident, h)), incident);
        double pdf = ((b.e + 1.0) * std::pow(ct, b.e)) / (2.0 * PI_ * 4.0 * std::fabs(dot(incident, h)));
        double f[3] = {0.0, 0.0, 0.0};
        if (incident.z * ex.z < 0.0) {
          pdf = 0.0;
        } else {
          double sc, fr;
          if (b.micro_eval(incident, ex, sc, fr))
            for (int c = 0; c < 3; ++c) f[c] = (b.ks[c] * fr) * sc;
        }
        if (pdf > 0.0) {
          const double w = std::fabs(ex.z) / pdf;
          for (int c = 0; c < 3; ++c) m[c] += w * f[c];
        }
      }
    for (int c = 0; c < 3; ++c) {
      m[c] = m[c] * (1.0 / 25.0);
      m[c] = smin(1.0, smax(0.0, m[c]));
      r[c] = r[c] + m[c];
    }
  }
  for (int c = 0; c < 3; ++c) ts[c] = smin(1.0, smax(0.0, r[c]));
}

struct FgSpec {
  const orc_accel *accel;
  const DLScene *s;
  const ParamView *v;
  const orc_irr_map *irr;
  const orc_photon_map *indirect;
  const std::vector<uint16_t> *dir_by_id;  // indirect photons' incident-direction codes by caller index
  const double *cdir;                      // 65536 x 3 decode table
  const std::vector<float> *irr_ext, *irr_int;
  double irr_max_dist, lookup_r;
  const double *cam;
  uint32_t width, height, spp, gs, gs_sqrt;
  uint64_t seed;
};

static void make_bsdf(const DLScene &s, const ParamView &v, uint32_t pr, V3 ng, V3 ns, Bsdf &b) {
  V3 t;
  if (std::fabs(ns.x) > std::fabs(ns.y)) {
    double il = 1.0 / std::sqrt(ns.x * ns.x + ns.z * ns.z);
    t = V3{-ns.z * il, 0.0, ns.x * il};
  } else {
    double il = 1.0 / std::sqrt(ns.y * ns.y + ns.z * ns.z);
    t = V3{0.0, ns.z * il, -ns.y * il};
  }
  V3 t2 = cross(ns, t);
  V3 tangent;
  if (len_sqr(t2) > DBL_EPS_) {
    tangent = cross(t2, ns);
    normalize(tangent);
  } else {
    tangent = t;
  }
  b.n = ns;
  b.ng = ng;
  b.e1 = tangent;
  b.e2 = cross(ns, tangent);
  b.plastic = s.mat[pr] == 1;
  for (int c = 0; c < 3; ++c) {
    b.kd[c] = v.kd[3 * pr + c];
    b.ks[c] = smin(1.0, smax(0.0, s.ks[3 * pr + c]));
  }
  b.e = v.e[pr];
}

// PhotonLTEIntegrator::_FinalGather (PhotonLTEIntegrator.cpp:387-577) at the camera hit of one
// sample; the stratified sequences come from fg_rnd. Returns the gather rays cast.
static uint64_t fg_sample(const FgSpec &F, uint64_t pix, uint32_t si, double L[3]) {
  const DLScene &s = *F.s;
  const double *cam = F.cam;
  const uint32_t px = (uint32_t)(pix % F.width), py = (uint32_t)(pix / F.width);
  const uint64_t sidx = pix * F.spp + si;
  double ix = (double)px + rnd(F.seed, sidx, 0), iy = (double)py + rnd(F.seed, sidx, 1);
  double cx = 2.0 * cam[12] * (ix - F.width * 0.5) / F.width;
  double cy = 2.0 * cam[13] * (iy - F.height * 0.5) / F.height;
  V3 d{cam[3] * cx + cam[4] * cy + cam[5], cam[6] * cx + cam[7] * cy + cam[8], cam[9] * cx + cam[10] * cy + cam[11]};
  normalize(d);
  double ray[8] = {cam[0], cam[1], cam[2], d.x, d.y, d.z, 0.0, DBL_INF_};
  orc_hit h;
  orc_hit_geom g;
  orc_accel_intersect_one(F.accel, ray, &h, &g);
  for (int c = 0; c < 3; ++c) L[c] = 0.0;
  if (h.prim == 0xffffffffu || F.gs == 0) return 0;
  const V3 wo = mul(d, -1.0);
  const V3 p{g.p[0], g.p[1], g.p[2]}, ng{g.ng[0], g.ng[1], g.ng[2]}, ns{g.ns[0], g.ns[1], g.ns[2]};
  Bsdf b;
  make_bsdf(s, *F.v, h.prim, ng, ns, b);
  const uint32_t gs = F.gs, gq = F.gs_sqrt;
  const double inv_n = 1.0 / gs, inv_q = 1.0 / gq;
  // StratifiedSampling1D / 2D with jitter (SamplingRoutines.h:327-373): BSDF 1D [0, gs), BSDF 2D
  // [gs, 3 gs), direction 1D [3 gs, 4 gs), direction 2D [4 gs, 6 gs).
  auto s1 = [&](uint32_t base, uint32_t i) { return (i + fg_rnd(F.seed, sidx, base + i)) * inv_n; };
  auto s2 = [&](uint32_t base, uint32_t i, double &u0, double &u1) {
    const uint32_t x = i % gq, y = i / gq;
    u0 = (x + fg_rnd(F.seed, sidx, base + 2 * i)) * inv_q;
    u1 = (y + fg_rnd(F.seed, sidx, base + 2 * i + 1)) * inv_q;
  };
  // the 32 nearest indirect photons with PhotonFilter (array order)
  uint32_t found = 0;
  uint32_t ids[32];
  if (F.indirect && orc_photon_map_size(F.indirect) > 0) {
    double pp[3] = {p.x, p.y, p.z}, nn[3] = {ng.x, ng.y, ng.z};
    float e0[3], e1[3];
    double d2[32];
    orc_photon_irradiance(F.indirect, 1, pp, nn, 32, F.lookup_r, 1, e0, e1, ids, d2, &found, 1);
  }
  const double inv_found = found > 0 ? 1.0 / found : 0.0;
  V3 pdir[32];
  for (uint32_t j = 0; j < found; ++j) {
    const double *dv = F.cdir + 3 * (size_t)(*F.dir_by_id)[ids[j]];
    pdir[j] = V3{dv[0], dv[1], dv[2]};
  }
  const double cone_pdf = 1.0 / (2.0 * PI_ * (1.0 - FG_COS_GATHER));
  auto photon_pdf_of = [&](V3 ex) {
    double pp = 0.0;
    for (uint32_t j = 0; j < found; ++j)
      if (dot(pdir[j], ex) > FG_COS_GATHER) pp += cone_pdf;
    return pp * inv_found;
  };
  std::vector<V3> gdir;
  std::vector<std::array<double, 3>> gw;
  for (uint32_t i = 0; i < gs; ++i) {
    const double uc = s1(0, i);
    double u0, u1;
    s2(gs, i, u0, u1);
    V3 ex, wo_l, wi_l;
    double bp, f[3];
    b.sample(wo, u0, u1, uc, ex, bp, f, wo_l, wi_l);
    if (bp > 0.0 && !(f[0] == 0.0 && f[1] == 0.0 && f[2] == 0.0)) {
      const double ppdf = photon_pdf_of(ex);
      const double w = power_heuristic((double)gs, bp, (double)gs, ppdf);
      const double k = w * std::fabs(dot(ex, ns)) / bp;
      gdir.push_back(ex);
      gw.push_back({f[0] * k, f[1] * k, f[2] * k});
    }
  }
  for (uint32_t i = 0; i < gs && found > 0; ++i) {
    const size_t si2 = (size_t)(s1(3 * gs, i) * found);
    double u0, u1;
    s2(4 * gs, i, u0, u1);
    V3 e2, e3;
    coordinate_system(pdir[si2], e2, e3);
    const double ct = (1.0 - u0) + u0 * FG_COS_GATHER;
    const double st = std::sqrt(1.0 - ct * ct);
    const double phi = u1 * 2.0 * PI_;
    const V3 loc{st * std::cos(phi), st * std::sin(phi), ct};
    const V3 ex = add(add(mul(e2, loc.x), mul(e3, loc.y)), mul(pdir[si2], loc.z));
    double f[3];
    b.evaluate(wo, ex, f);
    if (!(f[0] == 0.0 && f[1] == 0.0 && f[2] == 0.0)) {
      const double bp = b.pdf(wo, ex);
      const double ppdf = photon_pdf_of(ex);
      const double w = power_heuristic((double)gs, ppdf, (double)gs, bp);
      const double k = w * std::fabs(dot(ex, ns)) / ppdf;
      gdir.push_back(ex);
      gw.push_back({f[0] * k, f[1] * k, f[2] * k});
    }
  }
  double rad[3] = {0.0, 0.0, 0.0};
  for (size_t i = 0; i < gdir.size(); ++i) {
    const V3 dir = gdir[i];
    const double mint = orc_next_min_t(F.accel, h.prim, &g, &dir.x);
    double gr[8] = {p.x, p.y, p.z, dir.x, dir.y, dir.z, mint, DBL_INF_};
    orc_hit h2;
    orc_hit_geom g2;
    orc_accel_intersect_one(F.accel, gr, &h2, &g2);
    if (h2.prim == 0xffffffffu) continue;
    const V3 gd = mul(dir, -1.0);
    const V3 ng2{g2.ng[0], g2.ng[1], g2.ng[2]}, ns2{g2.ns[0], g2.ns[1], g2.ns[2]};
    double qp[3] = {g2.p[0], g2.p[1], g2.p[2]}, qn[3] = {ng2.x, ng2.y, ng2.z};
    uint32_t id = 0xffffffffu;
    orc_irr_map_nearest(F.irr, 1, qp, qn, F.irr_max_dist, &id, 1);
    if (id == 0xffffffffu) continue;
    Bsdf b2;
    make_bsdf(s, *F.v, h2.prim, ng2, ns2, b2);
    double ts[3];
    total_scattering(b2, gd, ts);
    const bool ext = dot(gd, ng2) > 0.0;
    const float *es = ext ? &(*F.irr_ext)[3 * id] : &(*F.irr_int)[3 * id];
    const float *eo = ext ? &(*F.irr_int)[3 * id] : &(*F.irr_ext)[3 * id];
    for (int c = 0; c < 3; ++c) {
      double t = 0.0;
      t = t + ts[c] * (double)es[c];
      t = t + 0.0 * (double)eo[c];  // no transmission lobe: TotalScattering of BSDF_ALL_TRANSMISSION is 0
      t = t * INV_PI_;
      rad[c] = rad[c] + (t * 1.0) * gw[i][c];
    }
  }
  const double inv = 1.0 / (double)gs;
  for (int c = 0; c < 3; ++c) L[c] = rad[c] * inv;
  return gdir.size();
}

// ------------------------------------------------------------------ photon shooting

// SamplingRoutines::RadicalInverse (SamplingRoutines.h:492-508).
static double radical_inverse(uint32_t n, uint32_t base) {
  double ret = 0.0;
  double inv_base = 1.0 / base, inv = inv_base;
  while (n > 0) {
    int d = (int)(n % base);
    ret += d * inv;
    n /= base;
    inv *= inv_base;
  }
  return ret;
}

// Counter-based stream of a photon path's later-bounce BSDF samples and Russian-roulette draws (the
// reference draws them from a per-chunk RandomGenerator): dims 4k..4k+2 = BSDF sample of
// intersection k >= 2, 4k+3 = roulette at intersection k; at most 1023 intersections per path.
static inline double shoot_rnd(uint64_t seed, uint64_t path, uint32_t dim) {
  uint64_t h = mix64((seed ^ 0xBB67AE8584CAA73Bull) ^ mix64(path * 4096ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}
static const uint32_t SHOOT_MAX_ISECT = 1023;

struct ShotPhoton {
  float p[3], w[3];
  uint16_t dir, normal;
};

// PhotonLTEIntegrator::_GetLightsPowerCDF (PhotonLTEIntegrator.cpp:80-124) over area lights:
// luminance of Power() = (pi * area) * radiance, accumulated, divided by the total.
static void lights_power_cdf(const DLScene &s, std::vector<double> &cdf) {
  const size_t nl = s.lights.size();
  cdf.assign(nl, 0.0);
  if (nl == 0) return;
  for (size_t l = 0; l < nl; ++l) {
    double power[3];
    for (int c = 0; c < 3; ++c) power[c] = (PI_ * s.lights[l].area) * s.lights[l].radiance[c];
    cdf[l] = luminance(power);
  }
  for (size_t i = 1; i < nl; ++i) cdf[i] += cdf[i - 1];
  const double total = cdf[nl - 1];
  if (total > DBL_EPS_)
    for (size_t i = 0; i < nl; ++i) cdf[i] /= total;
  else
    cdf[nl - 1] = 1.0;
}

struct ShootSpec {
  const orc_accel *accel;
  const DLScene *s;
  const ParamView *v;
  std::vector<double> light_cdf;
  uint64_t seed;
};

// One photon path of PhotonsShootingFilter::operator() (PhotonShootingPipeline.cpp:143-243) with
// DiffuseAreaLightSource::SamplePhoton (DiffuseAreaLightSource.cpp:118-150). Photons go to out[0]
// caustic, out[1] direct, out[2] indirect unless that map is done.
static void shoot_path(const ShootSpec &S, uint64_t path, const bool done[3], std::vector<ShotPhoton> out[3]) {
  const DLScene &s = *S.s;
  const uint32_t n1 = (uint32_t)path + 1u;
  const double ps0 = radical_inverse(n1, 2), ps1 = radical_inverse(n1, 3);
  const double ds0 = radical_inverse(n1, 5), ds1 = radical_inverse(n1, 7);
  double light_pdf = 0.0;
  const size_t li = search_cdf(S.light_cdf.data(), S.light_cdf.size(), radical_inverse(n1, 11), &light_pdf);
  const Light &L = s.lights[li];
  if (L.cdf.empty()) return;  // pdf 0
  const size_t ti = search_cdf(L.cdf.data(), L.cdf.size(), radical_inverse(n1, 13), nullptr);
  const V3 ln = L.normal[ti];
  const double tmp = std::sqrt(ps0), b1 = 1.0 - tmp, b2 = ps1 * tmp;
  const V3 sp = add(add(mul(L.v0[ti], b1), mul(L.v1[ti], b2)), mul(L.v2[ti], 1.0 - b1 - b2));
  V3 e2, e3;
  coordinate_system(ln, e2, e3);
  double dx, dy;
  concentric_disk(ds0, ds1, dx, dy);
  const double dz = std::sqrt(smax(0.0, 1.0 - dx * dx - dy * dy));
  V3 dir = add(add(mul(e2, dx), mul(e3, dy)), mul(ln, dz));
  const double photon_pdf = (dz * INV_PI_) * (1.0 / L.area);
  const double cosl = dot(dir, ln);
  double w[3];
  for (int c = 0; c < 3; ++c) w[c] = L.radiance[c] * cosl;
  if (photon_pdf == 0.0 || (w[0] == 0.0 && w[1] == 0.0 && w[2] == 0.0)) return;
  const double inv = 1.0 / (photon_pdf * light_pdf);
  for (int c = 0; c < 3; ++c) w[c] = w[c] * inv;
  bool specular_path = true;
  uint32_t n_isect = 0;
  double ray[8] = {sp.x, sp.y, sp.z, dir.x, dir.y, dir.z, 1e-4, DBL_INF_};
  while (true) {
    orc_hit h;
    orc_hit_geom g;
    orc_accel_intersect_one(S.accel, ray, &h, &g);
    if (h.prim == 0xffffffffu) break;
    ++n_isect;
    const V3 d{ray[3], ray[4], ray[5]};
    const V3 incident = mul(d, -1.0);
    const V3 p{g.p[0], g.p[1], g.p[2]}, ng{g.ng[0], g.ng[1], g.ng[2]}, ns{g.ns[0], g.ns[1], g.ns[2]};
    Bsdf b;
    make_bsdf(s, *S.v, h.prim, ng, ns, b);
    // Deposit (matte and plastic surfaces always have a non-specular component).
    ShotPhoton ph;
    for (int c = 0; c < 3; ++c) {
      ph.p[c] = (float)g.p[c];
      ph.w[c] = (float)w[c];
    }
    ph.dir = orc_cdir_encode(incident.x, incident.y, incident.z);
    ph.normal = orc_cdir_encode(ng.x, ng.y, ng.z);
    if (n_isect == 1) {
      if (!done[1]) out[1].push_back(ph);
    } else if (specular_path) {
      if (!done[0]) out[0].push_back(ph);
    } else {
      if (!done[2]) out[2].push_back(ph);
    }
    double u0, u1, uc;
    if (n_isect == 1) {
      u0 = radical_inverse(n1, 17);
      u1 = radical_inverse(n1, 19);
      uc = radical_inverse(n1, 23);
    } else {
      u0 = shoot_rnd(S.seed, path, 4 * n_isect);
      u1 = shoot_rnd(S.seed, path, 4 * n_isect + 1);
      uc = shoot_rnd(S.seed, path, 4 * n_isect + 2);
    }
    V3 ex, wo_l, wi_l;
    double bp, f[3];
    b.sample(incident, u0, u1, uc, ex, bp, f, wo_l, wi_l);
    if (bp == 0.0) break;
    double wn[3];
    const double inv_bp = 1.0 / bp;
    for (int c = 0; c < 3; ++c) wn[c] = (w[c] * f[c]) * inv_bp;
    const double cs = std::fabs(dot(ex, b.n));  // no specular lobes: always the cosine
    for (int c = 0; c < 3; ++c) wn[c] = wn[c] * cs;
    const double cp = smin(1.0, luminance(wn) / luminance(w));
    if (shoot_rnd(S.seed, path, 4 * n_isect + 3) > cp) break;
    const double inv_cp = 1.0 / cp;
    for (int c = 0; c < 3; ++c) w[c] = wn[c] * inv_cp;
    specular_path = false;  // the sampled component is never specular here
    if (done[2]) break;
    if (n_isect >= SHOOT_MAX_ISECT) break;
    const double mint = orc_next_min_t(S.accel, h.prim, &g, &ex.x);
    const double r2[8] = {p.x, p.y, p.z, ex.x, ex.y, ex.z, mint, DBL_INF_};
    std::memcpy(ray, r2, sizeof(ray));
  }
}

}  // namespace orc

using namespace orc;

struct orc_dl_scene {
  DLScene s;
  ParamView nominal;  // the scene's own materials
};

extern "C" {

orc_dl_scene *orc_dl_scene_create(const orc_accel *accel, const uint8_t *mat_type, const double *kd, const double *ks,
                                  const double *roughness, uint32_t n_lights, const uint32_t *light_prim,
                                  const double *light_radiance) {
  return orc_dl_scene_create_groups(accel, mat_type, kd, ks, roughness, n_lights, light_prim, light_radiance, 0,
                                    nullptr);
}

orc_dl_scene *orc_dl_scene_create_groups(const orc_accel *accel, const uint8_t *mat_type, const double *kd,
                                         const double *ks, const double *roughness, uint32_t n_lights,
                                         const uint32_t *light_prim, const double *light_radiance, uint32_t n_groups,
                                         const uint32_t *prim_group) {
  orc_dl_scene *h = new orc_dl_scene();
  DLScene &s = h->s;
  uint32_t np = orc_accel_prim_count(accel);
  s.mat.assign(mat_type, mat_type + np);
  s.kd.assign(kd, kd + 3 * (size_t)np);
  s.ks.assign(ks, ks + 3 * (size_t)np);
  s.rough.assign(roughness, roughness + np);
  s.light_of_prim.assign(np, -1);
  s.n_groups = n_groups;
  s.group.assign(np, NO_GROUP);
  if (n_groups > 0 && prim_group) s.group.assign(prim_group, prim_group + np);
  for (uint32_t l = 0; l < n_lights; ++l) {
    Light L;
    L.prim = light_prim[l];
    for (int c = 0; c < 3; ++c) L.radiance[c] = light_radiance[3 * l + c];
    s.light_of_prim[L.prim] = (int)l;
    uint32_t nt = orc_accel_mesh_tri_count(accel, L.prim);
    bool inv = orc_accel_mesh_inverted(accel, L.prim) != 0;
    L.cdf.assign(nt, 0.0);
    BBoxF bf;
    for (uint32_t t = 0; t < nt; ++t) {
      float v[9];
      orc_accel_mesh_tri(accel, L.prim, t, v);
      F3 a{v[0], v[1], v[2]}, b{v[3], v[4], v[5]}, c{v[6], v[7], v[8]};
      L.cdf[t] = tri_area_f(a, b, c);
      if (t > 0) L.cdf[t] += L.cdf[t - 1];
      L.normal.push_back(tri_normal_f(a, b, c, inv));
      L.v0.push_back(to_d(a));
      L.v1.push_back(to_d(b));
      L.v2.push_back(to_d(c));
      bf.unite(a);
      bf.unite(b);
      bf.unite(c);
    }
    if (nt > 0) {
      L.area = L.cdf.back();
      if (L.area > 0.0) {
        double inv_area = 1.0 / L.area;
        for (double &x : L.cdf) x *= inv_area;
      }
    }
    L.bbox.unite(bf);
    // Power = (pi * area) * radiance; intensity = Luminance(power) / (pi * bbox area)
    double power[3];
    for (int c = 0; c < 3; ++c) power[c] = (PI_ * L.area) * L.radiance[c];
    L.intensity = luminance(power) / (PI_ * L.bbox.area());
    s.lights.push_back(L);
  }
  make_view(s, nullptr, h->nominal);
  return h;
}

void orc_dl_scene_destroy(orc_dl_scene *h) { delete h; }

// ---- single-function hooks for the oracle's own checks (tests only) ----

// A BSDF in its local frame (normal = z, geometric normal = shading normal).
static Bsdf local_bsdf(int plastic, const double *kd, const double *ks, double e) {
  Bsdf b;
  b.n = V3{0.0, 0.0, 1.0};
  b.ng = b.n;
  b.e1 = V3{1.0, 0.0, 0.0};
  b.e2 = V3{0.0, 1.0, 0.0};
  b.plastic = plastic != 0;
  for (int c = 0; c < 3; ++c) {
    b.kd[c] = kd ? kd[c] : 0.0;
    b.ks[c] = ks ? ks[c] : 0.0;
  }
  b.e = e;
  return b;
}

double orc_blinn_d(double e, const double *h) {
  Bsdf b = local_bsdf(1, nullptr, nullptr, e);
  return b.blinn_d(V3{h[0], h[1], h[2]});
}

double orc_fresnel_dielectric(double cos_i) { return fresnel_dielectric(cos_i); }

int orc_microfacet_eval(double e, const double *wo, const double *wi, double *scalar, double *fres) {
  Bsdf b = local_bsdf(1, nullptr, nullptr, e);
  return b.micro_eval(V3{wo[0], wo[1], wo[2]}, V3{wi[0], wi[1], wi[2]}, *scalar, *fres) ? 1 : 0;
}

double orc_microfacet_pdf(double e, const double *wo, const double *wi) {
  Bsdf b = local_bsdf(1, nullptr, nullptr, e);
  return b.micro_pdf(V3{wo[0], wo[1], wo[2]}, V3{wi[0], wi[1], wi[2]});
}

void orc_bsdf_local_eval(int plastic, const double *kd, const double *ks, double e, const double *wo,
                         const double *wi, double *f, double *pdf) {
  Bsdf b = local_bsdf(plastic, kd, ks, e);
  V3 o{wo[0], wo[1], wo[2]}, i{wi[0], wi[1], wi[2]};
  b.evaluate(o, i, f);
  *pdf = b.pdf(o, i);
}

double orc_bsdf_local_sample(int plastic, const double *kd, const double *ks, double e, const double *wo, double u0,
                             double u1, double uc, double *wi, double *f) {
  Bsdf b = local_bsdf(plastic, kd, ks, e);
  V3 w, wo_l, wi_l;
  double p = 0.0;
  b.sample(V3{wo[0], wo[1], wo[2]}, u0, u1, uc, w, p, f, wo_l, wi_l);
  wi[0] = w.x;
  wi[1] = w.y;
  wi[2] = w.z;
  return p;
}

// IrradianceLightsSamplingStrategy::GetLightsCDF(point, normal) with the scene's own lights.
void orc_dl_lights_cdf(const orc_dl_scene *h, const double *p, const double *n, double *cdf) {
  if (h->s.lights.empty()) return;
  lights_cdf(h->s, h->nominal.intensity.data(), V3{p[0], p[1], p[2]}, V3{n[0], n[1], n[2]}, cdf);
}

// DiffuseAreaLightSource::SampleLighting of light l from point p: returns the triangle, writes the
// direction, distance, solid-angle pdf and whether the lit side faces p.
uint32_t orc_dl_light_sample(const orc_dl_scene *h, uint32_t l, const double *p, double tri_u, double u0, double u1,
                             double *wi, double *dist, double *pdf, int *lit) {
  size_t ti;
  V3 w;
  bool lt;
  light_sample(h->s.lights[l], V3{p[0], p[1], p[2]}, tri_u, u0, u1, ti, w, *dist, *pdf, lt);
  wi[0] = w.x;
  wi[1] = w.y;
  wi[2] = w.z;
  *lit = lt ? 1 : 0;
  return (uint32_t)ti;
}

double orc_dl_light_pdf(const orc_dl_scene *h, uint32_t l, uint32_t tri, const double *wi, double dist) {
  return light_pdf(h->s.lights[l], tri, V3{wi[0], wi[1], wi[2]}, dist);
}

// Area (sum of float triangle areas), Power luminance weight (pi * area) and intensity of light l.
void orc_dl_light_info(const orc_dl_scene *h, uint32_t l, double *area, double *intensity) {
  *area = h->s.lights[l].area;
  *intensity = h->s.lights[l].intensity;
}

uint32_t orc_dl_param_count(const orc_dl_scene *h) { return 2 * h->s.n_groups + (uint32_t)h->s.lights.size(); }

// Renders pixels [p0, p1) of a width x height image at spp samples each.
// cam: origin (3), camera-to-world rotation (9, row-major), x_tan, y_tan. Output rgb (float, 3 per pixel of
// the range), and counters: [0] primary rays, [1] shadow rays, [2] BSDF rays.
void orc_dl_render(const orc_accel *accel, const orc_dl_scene *dls, const double *cam, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples, uint64_t seed, uint64_t p0,
                   uint64_t p1, float *rgb, uint64_t *counters, uint32_t *primary_prim, uint32_t *primary_tri,
                   int nthreads) {
  FrameSpec F{accel, &dls->s, cam, width, height, spp, n_light_samples, n_bsdf_samples, seed, &dls->nominal,
              &dls->nominal};
  uint64_t c_prim = 0, c_shadow = 0, c_bsdf = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : c_prim, c_shadow, c_bsdf)
  for (int64_t pix = (int64_t)p0; pix < (int64_t)p1; ++pix) {
    double acc[3] = {0, 0, 0};
    for (uint32_t si = 0; si < spp; ++si) {
      double L[3];
      uint32_t hp, ht;
      bool good = dl_sample<double>(F, (uint64_t)pix, si, L, hp, ht, c_shadow, c_bsdf);
      ++c_prim;
      if (primary_prim) {
        primary_prim[(pix - p0) * spp + si] = hp;
        primary_tri[(pix - p0) * spp + si] = ht;
      }
      if (good)
        for (int c = 0; c < 3; ++c) acc[c] += L[c];
    }
    for (int c = 0; c < 3; ++c) rgb[3 * (pix - p0) + c] = (float)(acc[c] / (double)spp);
  }
  if (counters) {
    counters[0] = c_prim;
    counters[1] = c_shadow;
    counters[2] = c_bsdf;
  }
}

// The replayed estimator F(theta_eval; theta_sample) over pixels [p0, p1): mean radiance per pixel
// in double.
void orc_dl_render_replay(const orc_accel *accel, const orc_dl_scene *dls, const double *theta_eval,
                          const double *theta_sample, const double *cam, uint32_t width, uint32_t height, uint32_t spp,
                          uint32_t n_light_samples, uint32_t n_bsdf_samples, uint64_t seed, uint64_t p0, uint64_t p1,
                          double *img, int nthreads) {
  ParamView vs, ve;
  make_view(dls->s, theta_sample, vs);
  make_view(dls->s, theta_eval, ve);
  FrameSpec F{accel, &dls->s, cam, width, height, spp, n_light_samples, n_bsdf_samples, seed, &vs, &ve};
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t pix = (int64_t)p0; pix < (int64_t)p1; ++pix) {
    double acc[3] = {0, 0, 0};
    uint64_t cs = 0, cb = 0;
    for (uint32_t si = 0; si < spp; ++si) {
      double L[3];
      uint32_t hp, ht;
      if (dl_sample<double>(F, (uint64_t)pix, si, L, hp, ht, cs, cb))
        for (int c = 0; c < 3; ++c) acc[c] += L[c];
    }
    for (int c = 0; c < 3; ++c) img[3 * (pix - p0) + c] = acc[c] / (double)spp;
  }
}

// Gradient of sum_{pixel, c} adj[pixel, c] * image[pixel, c] over pixels [p0, p1) with respect to
// theta (theta_eval = theta_sample = theta), by forward-mode dual numbers per sample; img (optional)
// receives the image.
void orc_dl_gradient(const orc_accel *accel, const orc_dl_scene *dls, const double *theta, const double *cam,
                     uint32_t width, uint32_t height, uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples,
                     uint64_t seed, uint64_t p0, uint64_t p1, const double *adj, double *grad, double *img,
                     int nthreads) {
  const DLScene &s = dls->s;
  ParamView v;
  make_view(s, theta, v);
  FrameSpec F{accel, &s, cam, width, height, spp, n_light_samples, n_bsdf_samples, seed, &v, &v};
  const uint32_t G = s.n_groups, nl = (uint32_t)s.lights.size(), n_params = 2 * G + nl;
  std::vector<double> total(n_params, 0.0);
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<double> part(n_params, 0.0);
#pragma omp for schedule(dynamic, 16)
    for (int64_t pix = (int64_t)p0; pix < (int64_t)p1; ++pix) {
      double acc[3] = {0, 0, 0};
      uint64_t cs = 0, cb = 0;
      const double *a = adj + 3 * (pix - p0);
      for (uint32_t si = 0; si < spp; ++si) {
        Dual L[3];
        uint32_t hp, ht;
        if (!dl_sample<Dual>(F, (uint64_t)pix, si, L, hp, ht, cs, cb)) continue;
        for (int c = 0; c < 3; ++c) acc[c] += L[c].v;
        if (hp == 0xffffffffu) continue;
        const double inv_spp = 1.0 / (double)spp;
        double dj[ND];
        for (int j = 0; j < ND; ++j) {
          dj[j] = 0.0;
          for (int c = 0; c < 3; ++c) dj[j] += a[c] * L[c].d[j];
          dj[j] *= inv_spp;
        }
        if (s.group[hp] != NO_GROUP) {
          part[s.group[hp]] += dj[0];
          part[G + s.group[hp]] += dj[1];
        }
        for (uint32_t l = 0; l < nl; ++l) part[2 * G + l] += dj[2 + l];
      }
      if (img)
        for (int c = 0; c < 3; ++c) img[3 * (pix - p0) + c] = acc[c] / (double)spp;
    }
#pragma omp critical
    for (uint32_t j = 0; j < n_params; ++j) total[j] += part[j];
  }
  for (uint32_t j = 0; j < n_params; ++j) grad[j] = total[j];
}

}  // extern "C"

// The tables a final gather reads (direction decode table, indirect photons' direction codes by
// caller index, irradiance photons' estimates) and the FgSpec over them.
struct FgData {
  std::vector<double> table;
  std::vector<uint16_t> dir_by_id;
  std::vector<float> e_ext, e_int;
  FgSpec F;
};

static void init_fg(FgData &D, const orc_accel *accel, const orc_dl_scene *dls, const orc_irr_map *irr,
                    const orc_photon_map *indirect, double scene_area, const double *cam, uint32_t width,
                    uint32_t height, uint32_t spp, uint32_t gather_samples, uint64_t seed) {
  D.table.assign(3 * 65536, 0.0);
  orc_cdir_table(D.table.data());
  uint64_t n_ind = indirect ? orc_photon_map_size(indirect) : 0;
  if (n_ind > 0) {
    std::vector<float> pos(3 * n_ind);
    std::vector<uint16_t> dir(n_ind);
    std::vector<uint32_t> perm(n_ind);
    orc_photon_map_points(indirect, pos.data(), dir.data(), nullptr);
    orc_photon_map_dump(indirect, perm.data(), nullptr, nullptr);
    D.dir_by_id.assign(n_ind, 0);
    for (uint64_t i = 0; i < n_ind; ++i) D.dir_by_id[perm[i]] = dir[i];
  }
  uint64_t n_irr = 0;
  double irr_max = 0.0;
  orc_irr_map_info(irr, &n_irr, &irr_max);
  D.e_ext.assign(3 * n_irr, 0.0f);
  D.e_int.assign(3 * n_irr, 0.0f);
  orc_irr_map_export(irr, nullptr, nullptr, D.e_ext.data(), D.e_int.data());
  FgSpec &F = D.F;
  F.accel = accel;
  F.s = &dls->s;
  F.v = &dls->nominal;
  F.irr = irr;
  F.indirect = indirect;
  F.dir_by_id = &D.dir_by_id;
  F.cdir = D.table.data();
  F.irr_ext = &D.e_ext;
  F.irr_int = &D.e_int;
  F.irr_max_dist = irr_max;
  // PhotonLTEIntegrator.cpp:460-462: the radius holding 32 indirect photons on average
  F.lookup_r = n_ind > 0 ? std::sqrt((scene_area / (double)n_ind) * 32 * INV_PI_) : 0.0;
  F.cam = cam;
  F.width = width;
  F.height = height;
  F.spp = spp;
  F.gs_sqrt = (uint32_t)std::ceil(std::sqrt((double)gather_samples));
  F.gs = F.gs_sqrt * F.gs_sqrt;
  F.seed = seed;
}

extern "C" {

void orc_fg_render(const orc_accel *accel, const orc_dl_scene *dls, const orc_irr_map *irr,
                   const orc_photon_map *indirect, double scene_area, const double *cam, uint32_t width,
                   uint32_t height, uint32_t spp, uint32_t gather_samples, uint64_t seed, uint64_t p0, uint64_t p1,
                   double *radiance, uint64_t *gather_rays, int nthreads) {
  FgData D;
  init_fg(D, accel, dls, irr, indirect, scene_area, cam, width, height, spp, gather_samples, seed);
  const FgSpec &F = D.F;
  uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : rays)
  for (int64_t pix = (int64_t)p0; pix < (int64_t)p1; ++pix)
    for (uint32_t si = 0; si < spp; ++si) {
      double L[3];
      rays += fg_sample(F, (uint64_t)pix, si, L);
      for (int c = 0; c < 3; ++c) radiance[3 * (((uint64_t)pix - p0) * spp + si) + c] = L[c];
    }
  if (gather_rays) *gather_rays = rays;
}

// PhotonLTEIntegrator::_SurfaceRadiance (PhotonLTEIntegrator.cpp:320-385) for matte and plastic
// surfaces, one camera sample at a time: ((Le + ComputeDirectLighting) + _LookupCausticRadiance)
// + _FinalGather, then the renderer's NaN / negative-luminance / infinite rule on the sum
// (SamplerBasedRenderer.cpp:458-478) and the pixel mean. Matte and plastic surfaces scatter no
// photon specularly, so the caustic map holds no photons and its lookup adds a zero spectrum
// (PhotonLTEIntegrator.cpp:585-586). irr may be NULL (no final gather). The direct-lighting part
// draws from the direct-lighting stream, the final gather from its own.
void orc_pm_render(const orc_accel *accel, const orc_dl_scene *dls, const orc_irr_map *irr,
                   const orc_photon_map *indirect, double scene_area, const double *cam, uint32_t width,
                   uint32_t height, uint32_t spp, uint32_t n_light_samples, uint32_t n_bsdf_samples,
                   uint32_t gather_samples, uint64_t seed, uint64_t p0, uint64_t p1, float *rgb, uint64_t *counters,
                   int nthreads) {
  FrameSpec F{accel, &dls->s, cam, width, height, spp, n_light_samples, n_bsdf_samples, seed, &dls->nominal,
              &dls->nominal};
  FgData D;
  if (irr) init_fg(D, accel, dls, irr, indirect, scene_area, cam, width, height, spp, gather_samples, seed);
  uint64_t c_prim = 0, c_shadow = 0, c_bsdf = 0, c_gather = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1) reduction(+ : c_prim, c_shadow, c_bsdf, c_gather)
  for (int64_t pix = (int64_t)p0; pix < (int64_t)p1; ++pix) {
    double acc[3] = {0, 0, 0};
    for (uint32_t si = 0; si < spp; ++si) {
      double L[3], Lg[3] = {0.0, 0.0, 0.0};
      uint32_t hp, ht;
      (void)dl_sample<double>(F, (uint64_t)pix, si, L, hp, ht, c_shadow, c_bsdf);
      ++c_prim;
      if (irr) c_gather += fg_sample(D.F, (uint64_t)pix, si, Lg);
      double T[3];
      for (int c = 0; c < 3; ++c) T[c] = hp != 0xffffffffu ? (L[c] + 0.0) + Lg[c] : L[c];
      const bool bad = std::isnan(T[0]) || std::isnan(T[1]) || std::isnan(T[2]) || luminance(T) < -DBL_EPS_ ||
                       std::isinf(T[0]) || std::isinf(T[1]) || std::isinf(T[2]);
      if (!bad)
        for (int c = 0; c < 3; ++c) acc[c] += T[c];
    }
    for (int c = 0; c < 3; ++c) rgb[3 * (pix - p0) + c] = (float)(acc[c] / (double)spp);
  }
  if (counters) {
    counters[0] = c_prim;
    counters[1] = c_shadow;
    counters[2] = c_bsdf;
    counters[3] = c_gather;
  }
}

// PhotonLTEIntegrator::ShootPhotons (PhotonLTEIntegrator.cpp:247-288) with the pipeline's input
// and merging filters (PhotonShootingPipeline.cpp:48-92, 254-267) run one round of
// paths_per_round paths at a time, in path order: a map takes a round's photons (and counts its
// paths) unless it already held its maximum when the round began.
struct orc_photon_set {
  std::vector<ShotPhoton> m[3];
  uint64_t paths[3] = {0, 0, 0};
  uint64_t total_paths = 0;
};

orc_photon_set *orc_photon_shoot(const orc_accel *accel, const orc_dl_scene *dls, uint64_t max_paths,
                                 const uint64_t *max_photons, uint32_t paths_per_round, uint64_t seed, int nthreads) {
  orc_photon_set *out = new orc_photon_set();
  if (dls->s.lights.empty() || max_paths == 0 || paths_per_round == 0) return out;
  ShootSpec S;
  S.accel = accel;
  S.s = &dls->s;
  S.v = &dls->nominal;
  S.seed = seed;
  lights_power_cdf(dls->s, S.light_cdf);
  const int nt = nthreads > 0 ? nthreads : 1;
  while (true) {
    bool done[3];
    for (int j = 0; j < 3; ++j) done[j] = out->m[j].size() >= max_photons[j];
    if (out->total_paths >= max_paths || (done[0] && done[1] && done[2])) break;
    bool stuck = out->total_paths > 10000000;
    for (int j = 0; j < 3; ++j) stuck = stuck && (done[j] || out->m[j].empty());
    if (stuck) break;
    const uint64_t first = out->total_paths;
    const uint64_t n = std::min<uint64_t>(paths_per_round, max_paths - first);
    out->total_paths += n;
    // paths in blocks of 256, merged in block order
    const uint64_t nb = (n + 255) / 256;
    std::vector<std::array<std::vector<ShotPhoton>, 3>> part(nb);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t bi = 0; bi < (int64_t)nb; ++bi) {
      const uint64_t p0 = first + (uint64_t)bi * 256, p1 = std::min(first + n, p0 + 256);
      for (uint64_t p = p0; p < p1; ++p) shoot_path(S, p, done, part[bi].data());
    }
    for (int j = 0; j < 3; ++j) {
      if (done[j]) continue;
      out->paths[j] += n;
      for (auto &pb : part) out->m[j].insert(out->m[j].end(), pb[j].begin(), pb[j].end());
    }
  }
  return out;
}

void orc_photon_set_info(const orc_photon_set *s, uint64_t *counts, uint64_t *paths, uint64_t *total_paths) {
  for (int j = 0; j < 3; ++j) {
    if (counts) counts[j] = s->m[j].size();
    if (paths) paths[j] = s->paths[j];
  }
  if (total_paths) *total_paths = s->total_paths;
}

void orc_photon_set_export(const orc_photon_set *s, uint32_t which, float *pos, float *weight, uint16_t *dir,
                           uint16_t *normal) {
  const std::vector<ShotPhoton> &v = s->m[which];
  for (size_t i = 0; i < v.size(); ++i) {
    for (int c = 0; c < 3; ++c) {
      if (pos) pos[3 * i + c] = v[i].p[c];
      if (weight) weight[3 * i + c] = v[i].w[c];
    }
    if (dir) dir[i] = v[i].dir;
    if (normal) normal[i] = v[i].normal;
  }
}

void orc_photon_set_destroy(orc_photon_set *s) { delete s; }

double orc_radical_inverse(uint32_t n, uint32_t base) { return radical_inverse(n, base); }

}  // extern "C"
