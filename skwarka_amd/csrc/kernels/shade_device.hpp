// Device helpers shared by the shading kernels (dl.hip, shoot.hip): the sample stream, light
// tables, the matte / plastic BSDF with its sampling, the power heuristic, ray records and the
// block-aggregated append. Formulas follow the reference files cited at each helper and in dl.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "geom_device.hpp"
#include "kernels.hpp"

namespace skw {
namespace {

using namespace dev;

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kInv2Pi = 0.15915494309189533577;
constexpr double kEps = 2.220446049250313080847e-16;
constexpr int kMaxLights = 16;
// Y row of the reference's spectrum-to-XYZ matrix (SpectrumRoutines.cpp:20-62), i.e. its CIE Y
// table summed over the 360-494, 495-574 and 575-830 nm bands and divided by the table's total.
__constant__ double kLumW[3] = {0x1.b38293da589a6p-5, 0x1.2b993cc221566p-1, 0x1.725d3400723f6p-2};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double rnd(uint64_t seed, uint64_t sample, uint32_t dim) {
  uint64_t h = mix64(seed ^ mix64(sample * 64ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------------ lights

__device__ double solid_angle3(D3 d1, D3 d2, D3 d3) {
  double det = dot3(d1, cross3(d2, d3));
  double denom = 1 + dot3(d1, d2) + dot3(d1, d3) + dot3(d2, d3);
  if (denom > 0.0) return fabs(atan(det / denom) * 2.0);
  if (denom < 0.0) return 2.0 * (kPi - fabs(atan(det / denom)));
  return 2.0 * kPi;
}

// Solid angle of the light's box seen from p, from the unit directions toward its corners
// (dirs[i]: corner i takes bmin on axis k where bit k of i is clear, bmax where it is set).
__device__ double subtended_solid_angle(D3 p, const DevLight &L, const D3 *dirs) {
  bool inside = true;
  double sum = 0.0;
  if (p.x < L.bmin[0] - kEps) {
    sum += solid_angle3(dirs[0], dirs[2], dirs[6]) + solid_angle3(dirs[0], dirs[4], dirs[6]);
    inside = false;
  } else if (p.x > L.bmax[0] + kEps) {
    sum += solid_angle3(dirs[1], dirs[3], dirs[7]) + solid_angle3(dirs[1], dirs[5], dirs[7]);
    inside = false;
  }
  if (p.y < L.bmin[1] - kEps) {
    sum += solid_angle3(dirs[0], dirs[1], dirs[5]) + solid_angle3(dirs[0], dirs[4], dirs[5]);
    inside = false;
  } else if (p.y > L.bmax[1] + kEps) {
    sum += solid_angle3(dirs[2], dirs[3], dirs[7]) + solid_angle3(dirs[2], dirs[6], dirs[7]);
    inside = false;
  }
  if (p.z < L.bmin[2] - kEps) {
    sum += solid_angle3(dirs[0], dirs[1], dirs[3]) + solid_angle3(dirs[0], dirs[2], dirs[3]);
    inside = false;
  } else if (p.z > L.bmax[2] + kEps) {
    sum += solid_angle3(dirs[4], dirs[5], dirs[7]) + solid_angle3(dirs[4], dirs[6], dirs[7]);
    inside = false;
  }
  return inside ? 4.0 * kPi : sum;
}

// The corner directions are normalized once per light and serve both the cosine average and the
// solid angle (the same values the two loops would each compute, summed in the same order).
__device__ void lights_cdf(const DlArgs &a, D3 p, D3 n, double *cdf) {
  const uint32_t nl = a.n_lights;
  for (uint32_t i = 0; i < nl; ++i) {
    const DevLight &L = a.lights[i];
    bool inside = p.x >= L.bmin[0] && p.y >= L.bmin[1] && p.z >= L.bmin[2] && p.x <= L.bmax[0] && p.y <= L.bmax[1] &&
                  p.z <= L.bmax[2];
    if (inside) {
      cdf[i] = kPi * L.intensity;
    } else {
      D3 dirs[8];
      for (int c = 0; c < 8; ++c) {
        D3 v{(c & 1) ? L.bmax[0] : L.bmin[0], (c & 2) ? L.bmax[1] : L.bmin[1], (c & 4) ? L.bmax[2] : L.bmin[2]};
        D3 d = sub3(v, p);
        normalize3(d);
        dirs[c] = d;
      }
      // corners in the order bmax first on every axis (corner j of the average is dirs[j ^ 7])
      double avg = 0.0;
      for (int j = 0; j < 8; ++j) avg += dmax(0.0, dot3(dirs[j ^ 7], n));
      avg *= 0.125;
      if (avg > 0.0) {
        if (avg < 0.1) avg = 0.1;
        cdf[i] = subtended_solid_angle(p, L, dirs) * L.intensity * avg;
      } else {
        cdf[i] = 0.0;
      }
    }
  }
  for (uint32_t i = 1; i < nl; ++i) cdf[i] += cdf[i - 1];
  if (cdf[nl - 1] > 0.0) {
    double inv = 1.0 / cdf[nl - 1];
    for (uint32_t i = 0; i < nl; ++i) cdf[i] *= inv;
  } else {
    double inv = 1.0 / (double)nl;
    for (uint32_t i = 0; i < nl; ++i) cdf[i] = ((double)i + 1.0) * inv;
  }
}

template <class Get>
__device__ uint32_t search_cdf(Get get, uint32_t n, double u) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (get(mid) <= u) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------------ BSDF

__device__ double fresnel_dielectric(double c) {
  double ni = 1.5, no = 1.0;
  c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
  if (c < 0.0) {
    double t = ni;
    ni = no;
    no = t;
  }
  double sin_t = no / ni * sqrt(dmax(0.0, 1.0 - c * c));
  if (sin_t >= 1.0) return 1.0;
  double cos_t = sqrt(dmax(0.0, 1.0 - sin_t * sin_t));
  c = fabs(c);
  double rpar = ((ni * c) - (no * cos_t)) / ((ni * c) + (no * cos_t));
  double rper = ((no * c) - (ni * cos_t)) / ((no * c) + (ni * cos_t));
  return (rpar * rpar + rper * rper) * 0.5;
}

struct Bsdf {
  D3 n, ng, e1, e2;
  bool plastic;
  double kd[3], ks[3];
  double e;

  __device__ D3 to_local(D3 v) const { return D3{dot3(v, e1), dot3(v, e2), dot3(v, n)}; }
  __device__ D3 to_world(D3 v) const { return add3(add3(mul3(e1, v.x), mul3(e2, v.y)), mul3(n, v.z)); }
  __device__ double blinn_pdf(D3 wo, D3 wi) const {
    D3 h = add3(wo, wi);
    normalize3(h);
    double ct = fabs(h.z);
    return ((e + 1.0) * pow(ct, e)) / (2.0 * kPi * 4.0 * fabs(dot3(wo, h)));
  }
  // Microfacet::Evaluate's scalar D*G/(4 cos_i cos_o) and Fresnel term; with DERIV also
  // d scalar / d exponent, from d D / d e = D / (e + 2) + D ln(cos_h).
  template <bool DERIV>
  __device__ bool micro_eval(D3 wo, D3 wi, double &scalar, double &fres, double &dscalar) const {
    double ci = fabs(wo.z), co = fabs(wi.z);
    if (ci < kEps || co < kEps) return false;
    D3 h = add3(wo, wi);
    if (fabs(h.x) < kEps && fabs(h.y) < kEps && fabs(h.z) < kEps) return false;
    normalize3(h);
    fres = fresnel_dielectric(dot3(wi, h));
    double ch = fabs(h.z);
    double tmp = 2.0 * ch / fabs(dot3(wo, h));
    double g = dmin(1.0, dmin(ci * tmp, co * tmp));
    double pw = pow(ch, e);
    double d = (e + 2.0) * kInv2Pi * pw;
    scalar = d * g / (4.0 * ci * co);
    if (DERIV) {
      double dd = pw > 0.0 ? kInv2Pi * pw * (1.0 + (e + 2.0) * log(ch)) : 0.0;  // D = 0 at cos_h = 0
      dscalar = dd * g / (4.0 * ci * co);
    }
    return true;
  }
  __device__ double micro_pdf(D3 wo, D3 wi) const { return (wo.z * wi.z < 0.0) ? 0.0 : blinn_pdf(wo, wi); }
  __device__ static double lambert_pdf(D3 wo, D3 wi) { return (wo.z * wi.z > 0.0) ? fabs(wi.z) * kInvPi : 0.0; }
  // Sum of the reflection BxDFs for local vectors. With DERIV, dfr = d f[c] / d roughness / ks[c]
  // (the exponent is 1/roughness; de_dr carries its derivative, 0 where clamped).
  template <bool DERIV>
  __device__ void eval_reflect_local(D3 wo, D3 wi, double f[3], double de_dr, double &dfr) const {
    for (int c = 0; c < 3; ++c) f[c] = kd[c] * kInvPi;
    if (DERIV) dfr = 0.0;
    if (plastic) {
      double sc, fr, dsc = 0.0;
      if (micro_eval<DERIV>(wo, wi, sc, fr, dsc)) {
        for (int c = 0; c < 3; ++c) f[c] += (ks[c] * fr) * sc;
        if (DERIV) dfr = fr * dsc * de_dr;
      }
    }
  }
  __device__ void eval_reflect_local(D3 wo, D3 wi, double f[3]) const {
    double unused;
    eval_reflect_local<false>(wo, wi, f, 0.0, unused);
  }
  template <bool DERIV>
  __device__ void evaluate(D3 wo_w, D3 wi_w, double f[3], double de_dr, double &dfr) const {
    D3 wo = to_local(wo_w), wi = to_local(wi_w);
    if (dot3(wo_w, ng) * dot3(wi_w, ng) > 0) {
      eval_reflect_local<DERIV>(wo, wi, f, de_dr, dfr);
    } else {
      f[0] = f[1] = f[2] = 0.0;
      if (DERIV) dfr = 0.0;
    }
  }
  __device__ void evaluate(D3 wo_w, D3 wi_w, double f[3]) const {
    double unused;
    evaluate<false>(wo_w, wi_w, f, 0.0, unused);
  }
  __device__ double pdf(D3 wo_w, D3 wi_w) const {
    D3 wo = to_local(wo_w), wi = to_local(wi_w);
    double p = lambert_pdf(wo, wi);
    if (!plastic) return p;
    p += micro_pdf(wo, wi);
    return p / 2.0;
  }
  // BSDF::Sample; wo_l / wi_l return the local vectors f was evaluated with.
  __device__ void sample(D3 wo_w, double u0, double u1, double uc, D3 &wi_w, double &pdf_out, double f[3], D3 &wo_l,
                         D3 &wi_l) const {
    const uint64_t num = plastic ? 2 : 1;
    uint64_t comp = (uint64_t)(uc * (double)num);
    if (comp > num - 1) comp = num - 1;
    D3 wo = to_local(wo_w);
    D3 wi;
    double p;
    if (comp == 0) {
      // ConcentricDiskSampling + CosineHemisphereSampling
      double s0 = 2.0 * u0 - 1.0, s1 = 2.0 * u1 - 1.0;
      double dx = 0.0, dy = 0.0;
      if (!(s0 == 0.0 && s1 == 0.0)) {
        double r, th;
        if (s0 >= -s1) {
          if (s0 > s1) {
            r = s0;
            th = s1 > 0.0 ? s1 / r : 8.0 + s1 / r;
          } else {
            r = s1;
            th = 2.0 - s0 / r;
          }
        } else {
          if (s0 <= s1) {
            r = -s0;
            th = 4.0 - s1 / r;
          } else {
            r = -s1;
            th = 6.0 + s0 / r;
          }
        }
        th *= kPi / 4.0;
        dx = r * cos(th);
        dy = r * sin(th);
      }
      double z = sqrt(dmax(0.0, 1.0 - dx * dx - dy * dy));
      wi = D3{dx, dy, z};
      if (wo.z < 0.0) wi.z *= -1.0;
      p = lambert_pdf(wo, wi);
    } else {
      double ct = pow(u0, 1.0 / (e + 1.0));
      double st = sqrt(dmax(0.0, 1.0 - ct * ct));
      double phi = u1 * 2.0 * kPi;
      D3 h{st * cos(phi), st * sin(phi), ct};
      if (wo.z * h.z < 0.0) h.z *= -1.0;
      wi = sub3(mul3(h, 2.0 * dot3(wo, h)), wo);
      p = ((e + 1.0) * pow(ct, e)) / (2.0 * kPi * 4.0 * fabs(dot3(wo, h)));
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
    if (dot3(wo_w, ng) * dot3(wi_w, ng) > 0) eval_reflect_local(wo, wi, f);
    else f[0] = f[1] = f[2] = 0.0;
  }
};

__device__ __forceinline__ double power_heuristic(double n1, double p1, double n2, double p2) {
  double f = n1 * p1, g = n2 * p2;
  return (f * f) / (f * f + g * g);
}

__device__ __forceinline__ void put_ray(skw_ray *r, D3 o, D3 d, double tmin, double tmax) {
  double2 *q = reinterpret_cast<double2 *>(r);
  q[0] = make_double2(o.x, o.y);
  q[1] = make_double2(o.z, d.x);
  q[2] = make_double2(d.y, d.z);
  q[3] = make_double2(tmin, tmax);
}

// Block-aggregated append: lanes with `want` get consecutive slots of *counter, one atomicAdd
// per block of kSampleBlock threads. A single device-wide counter takes about one atomic per
// 11 ns, so per-wave appends would bound the sample kernels; per block they cost 8x less. Every
// thread of the block must call this (it has two barriers).
constexpr uint32_t kSampleBlock = 512;
__device__ __forceinline__ int32_t append_slot_block(uint64_t *counter, bool want) {
  __shared__ uint32_t wave_count[kSampleBlock / 64];
  __shared__ unsigned long long base_s;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t mask = __ballot(want);
  if (lane == 0) wave_count[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) total += wave_count[w];
    base_s = total ? atomicAdd(reinterpret_cast<unsigned long long *>(counter), (unsigned long long)total) : 0ull;
  }
  __syncthreads();
  uint64_t off = base_s;
  for (uint32_t w = 0; w < wave; ++w) off += wave_count[w];
  if (!want) return -1;
  return (int32_t)(off + (uint64_t)__popcll(mask & ((1ull << lane) - 1ull)));
}

// MathRoutines::CoordinateSystem (MathRoutines.h:298-313).
__device__ __forceinline__ void coordinate_system(D3 e1, D3 &e2, D3 &e3) {
  if (fabs(e1.x) > fabs(e1.y)) {
    const double inv = 1.0 / sqrt(e1.x * e1.x + e1.z * e1.z);
    e2 = D3{-e1.z * inv, 0.0, e1.x * inv};
  } else {
    const double inv = 1.0 / sqrt(e1.y * e1.y + e1.z * e1.z);
    e2 = D3{0.0, e1.z * inv, -e1.y * inv};
  }
  e3 = cross3(e1, e2);
}

// The BSDF frame of a mesh without UVs or tangents (TriangleMesh.cpp:520-540): tangent from the
// shading normal's CoordinateSystem, re-orthogonalised.
__device__ __forceinline__ void bsdf_frame(D3 ns, D3 &tangent, D3 &bitangent) {
  D3 t;
  if (fabs(ns.x) > fabs(ns.y)) {
    double il = 1.0 / sqrt(ns.x * ns.x + ns.z * ns.z);
    t = D3{-ns.z * il, 0.0, ns.x * il};
  } else {
    double il = 1.0 / sqrt(ns.y * ns.y + ns.z * ns.z);
    t = D3{0.0, ns.z * il, -ns.y * il};
  }
  D3 t2 = cross3(ns, t);
  if (len_sqr3(t2) > kEps) {
    tangent = cross3(t2, ns);
    normalize3(tangent);
  } else {
    tangent = t;
  }
  bitangent = cross3(ns, tangent);
}

}  // namespace
}  // namespace skw
