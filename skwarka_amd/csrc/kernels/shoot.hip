// Photon shooting on gfx950, one bounce per launch (a wavefront over a round of paths).
//
// Per path, PhotonsShootingFilter::operator() (Raytracer/LTEIntegrators/PhotonLTEIntegrator/
// PhotonShootingPipeline.cpp:143-243):
//   k_shoot_start: radical-inverse samples of path index + 1 (bases 2, 3 position; 5, 7 direction;
//     11 light choice by the lights' power CDF; 13 triangle), DiffuseAreaLightSource::SamplePhoton
//     (LightSources/DiffuseAreaLightSource.cpp:118-150): point by the light's area CDF, cosine
//     direction around the triangle's float normal, weight = radiance * cos / (pdf * light pdf).
//   closest hit (the scene's traversal kernel), then k_shoot_bounce: deposit a photon (direct at
//     the first intersection, caustic while the path is specular, indirect otherwise; matte and
//     plastic surfaces always have a non-specular component and never scatter specularly), sample
//     the BSDF (radical inverse bases 17, 19, 23 at the first intersection, a counter-based
//     stream later), weight *= f |cos| / pdf, Russian roulette with continue probability
//     min(1, Y(new) / Y(old)), next ray from CoreUtils::GetNextMinT.
// Photons and surviving paths are appended with one atomic per 512-thread block; the round's
// photons are put in (path, intersection) order afterwards by a stable sort on the path index.
#include <hip/hip_runtime.h>

#include "shade_device.hpp"

namespace skw {
namespace {

using namespace dev;

// SamplingRoutines::RadicalInverse (SamplingRoutines.h:492-508).
__device__ __forceinline__ double radical_inverse(uint32_t n, uint32_t base) {
  double ret = 0.0;
  const double inv_base = 1.0 / base;
  double inv = inv_base;
  while (n > 0) {
    const int d = (int)(n % base);
    ret += d * inv;
    n /= base;
    inv *= inv_base;
  }
  return ret;
}

// Counter-based stream of a path's later-bounce BSDF samples and roulette draws: dims 4k..4k+2 =
// BSDF sample at intersection k >= 2, 4k+3 = roulette at intersection k (k <= 1023).
__device__ __forceinline__ double shoot_rnd(uint64_t seed, uint64_t path, uint32_t dim) {
  const uint64_t h = mix64((seed ^ 0xBB67AE8584CAA73Bull) ^ mix64(path * 4096ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}
constexpr uint32_t kMaxIsect = 1023;

// CompressedDirection(const Vector3D_d&) (Math/CompressedDirection.h:92-124): octant bits 13-15,
// then the first two coordinates of the L1-normalised direction on a 127-step grid folded into a
// 64-wide triangle.
__device__ __forceinline__ uint32_t cdir_encode(D3 v) {
  uint32_t data = 0;
  if (v.x < 0) {
    data |= 1u << 13;
    v.x = -v.x;
  }
  if (v.y < 0) {
    data |= 1u << 14;
    v.y = -v.y;
  }
  if (v.z < 0) {
    data |= 1u << 15;
    v.z = -v.z;
  }
  const double scale = 126.999 / (v.x + v.y + v.z);
  uint32_t x = (uint32_t)(unsigned char)(v.x * scale);
  uint32_t y = (uint32_t)(unsigned char)(v.y * scale);
  if (y >= 64) {
    x = 127 - x;
    y = 127 - y;
  }
  return data | (x << 6) | y;
}

__device__ __forceinline__ double luminance3(const double s[3]) {
  return kLumW[0] * s[0] + kLumW[1] * s[1] + kLumW[2] * s[2];
}

__device__ __forceinline__ void put_state(ShootPath *st, const double w[3], uint32_t path, uint32_t n_isect) {
  ShootPath p;
  p.w[0] = w[0];
  p.w[1] = w[1];
  p.w[2] = w[2];
  p.path = path;
  p.n_isect = n_isect;
  *st = p;
}

__global__ __launch_bounds__(kSampleBlock) void k_shoot_start(ShootArgs a) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool want = false;
  D3 sp{0, 0, 0}, dir{0, 0, 0};
  double w[3] = {0.0, 0.0, 0.0};
  const uint64_t path = a.first + t;
  if (t < a.n) {
    const uint32_t n1 = (uint32_t)path + 1u;
    const double ps0 = radical_inverse(n1, 2), ps1 = radical_inverse(n1, 3);
    const double ds0 = radical_inverse(n1, 5), ds1 = radical_inverse(n1, 7);
    const double ul = radical_inverse(n1, 11);
    const uint32_t l = search_cdf([&](uint32_t j) { return a.light_cdf[j]; }, a.n_lights, ul);
    const double light_pdf = l == 0 ? a.light_cdf[0] : a.light_cdf[l] - a.light_cdf[l - 1];
    const DevLight &L = a.lights[l];
    if (L.tri_count > 0) {
      const DevLightTri *lt = a.ltris + L.tri_off;
      const uint32_t ti = search_cdf([&](uint32_t j) { return lt[j].cdf; }, L.tri_count, radical_inverse(n1, 13));
      const DevLightTri &T = lt[ti];
      const D3 ln{T.n[0], T.n[1], T.n[2]};
      const double tmp = sqrt(ps0), b1 = 1.0 - tmp, b2 = ps1 * tmp;
      const D3 v0{T.v0[0], T.v0[1], T.v0[2]}, v1{T.v1[0], T.v1[1], T.v1[2]}, v2{T.v2[0], T.v2[1], T.v2[2]};
      sp = add3(add3(mul3(v0, b1), mul3(v1, b2)), mul3(v2, 1.0 - b1 - b2));
      D3 e2, e3;
      coordinate_system(ln, e2, e3);
      // ConcentricDiskSampling + CosineHemisphereSampling (SamplingRoutines.h:197-291)
      const double s0 = 2.0 * ds0 - 1.0, s1 = 2.0 * ds1 - 1.0;
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
      const double dz = sqrt(dmax(0.0, 1.0 - dx * dx - dy * dy));
      dir = add3(add3(mul3(e2, dx), mul3(e3, dy)), mul3(ln, dz));
      const double photon_pdf = (dz * kInvPi) * (1.0 / L.area);
      const double cosl = dot3(dir, ln);
      for (int c = 0; c < 3; ++c) w[c] = L.radiance[c] * cosl;
      if (photon_pdf != 0.0 && !(w[0] == 0.0 && w[1] == 0.0 && w[2] == 0.0)) {
        const double inv = 1.0 / (photon_pdf * light_pdf);
        for (int c = 0; c < 3; ++c) w[c] = w[c] * inv;
        want = true;
      }
    }
  }
  const int32_t slot = append_slot_block(a.counts + 0, want);
  if (want) {
    put_ray(a.rays[0] + slot, sp, dir, 1e-4, __builtin_huge_val());
    put_state(a.state[0] + slot, w, (uint32_t)path, 0);
  }
}

__global__ __launch_bounds__(kSampleBlock) void k_shoot_bounce(ShootArgs a) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t cur = a.cur;
  const uint64_t n_active = a.counts[cur];  // the grid covers an upper bound of it
  bool alive = false, next = false;
  int map = -1;
  uint32_t codes = 0, key = 0, path = 0, n_isect = 0;
  float fp[3] = {0.0f, 0.0f, 0.0f}, fw[3] = {0.0f, 0.0f, 0.0f};
  D3 p{0, 0, 0}, ex{0, 0, 0};
  double w[3] = {0.0, 0.0, 0.0}, mint = 0.0;
  if (j < n_active) {
    const skw_hit h = a.hits[j];
    alive = h.prim != kMiss;
    if (alive) {
      const ShootPath st = a.state[cur][j];
      const skw_ray &r = a.rays[cur][j];
      const D3 o{r.o[0], r.o[1], r.o[2]}, d{r.d[0], r.d[1], r.d[2]};
      const HitGeom g = hit_geometry(a.geom, o, d, h.prim, h.tri);
      path = st.path;
      n_isect = st.n_isect + 1;
      for (int c = 0; c < 3; ++c) w[c] = st.w[c];
      const D3 incident = mul3(d, -1.0);
      Bsdf b;
      b.n = g.ns;
      b.ng = g.ng;
      bsdf_frame(g.ns, b.e1, b.e2);
      const DevMat &m = a.mats[h.prim];
      b.plastic = m.type == 1;
      for (int c = 0; c < 3; ++c) {
        b.kd[c] = m.kd[c];
        b.ks[c] = m.ks[c];
      }
      b.e = m.exponent;
      // deposit: direct at the first intersection, indirect afterwards (no specular scattering)
      map = n_isect == 1 ? 1 : 2;
      if (a.done[map]) map = -1;
      p = g.p;
      fp[0] = (float)p.x;
      fp[1] = (float)p.y;
      fp[2] = (float)p.z;
      for (int c = 0; c < 3; ++c) fw[c] = (float)w[c];
      codes = cdir_encode(incident) | (cdir_encode(g.ng) << 16);
      key = (uint32_t)((uint64_t)path - a.first);
      // scatter
      double u0, u1, uc;
      if (n_isect == 1) {
        const uint32_t n1 = path + 1u;
        u0 = radical_inverse(n1, 17);
        u1 = radical_inverse(n1, 19);
        uc = radical_inverse(n1, 23);
      } else {
        u0 = shoot_rnd(a.seed, path, 4 * n_isect);
        u1 = shoot_rnd(a.seed, path, 4 * n_isect + 1);
        uc = shoot_rnd(a.seed, path, 4 * n_isect + 2);
      }
      double bp, f[3];
      D3 wo_l, wi_l;
      b.sample(incident, u0, u1, uc, ex, bp, f, wo_l, wi_l);
      if (bp != 0.0) {
        double wn[3];
        const double inv_bp = 1.0 / bp;
        for (int c = 0; c < 3; ++c) wn[c] = (w[c] * f[c]) * inv_bp;
        const double cs = fabs(dot3(ex, b.n));
        for (int c = 0; c < 3; ++c) wn[c] = wn[c] * cs;
        const double cp = dmin(1.0, luminance3(wn) / luminance3(w));
        if (!(shoot_rnd(a.seed, path, 4 * n_isect + 3) > cp)) {
          const double inv_cp = 1.0 / cp;
          for (int c = 0; c < 3; ++c) w[c] = wn[c] * inv_cp;
          if (!a.done[2] && n_isect < kMaxIsect) {
            mint = next_min_t(a.geom.prims[h.prim], g, ex);
            next = true;
          }
        }
      }
    }
  }
  // Photon appends, one map per call (the helper's LDS is reused: a barrier between calls).
  for (int m = 0; m < 3; ++m) {
    const int32_t slot = append_slot_block(a.counts + 2 + m, map == m);
    if (map == m) {
      a.ph_pos[m][3 * (uint64_t)slot] = fp[0];
      a.ph_pos[m][3 * (uint64_t)slot + 1] = fp[1];
      a.ph_pos[m][3 * (uint64_t)slot + 2] = fp[2];
      a.ph_w[m][3 * (uint64_t)slot] = fw[0];
      a.ph_w[m][3 * (uint64_t)slot + 1] = fw[1];
      a.ph_w[m][3 * (uint64_t)slot + 2] = fw[2];
      a.ph_codes[m][slot] = codes;
      a.ph_key[m][slot] = key;
    }
    __syncthreads();
  }
  const uint32_t nx = cur ^ 1u;
  const int32_t slot = append_slot_block(a.counts + nx, next);
  if (next) {
    put_ray(a.rays[nx] + slot, p, ex, mint, __builtin_huge_val());
    put_state(a.state[nx] + slot, w, path, n_isect);
  }
}

__global__ __launch_bounds__(256) void k_photon_gather_records(const float *pos_in, const float *w_in,
                                                               const uint32_t *codes_in, const uint32_t *order,
                                                               uint64_t n, float *pos_out, float *w_out,
                                                               uint32_t *codes_out) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint64_t i = order[j];
  for (int c = 0; c < 3; ++c) {
    pos_out[3 * j + c] = pos_in[3 * i + c];
    w_out[3 * j + c] = w_in[3 * i + c];
  }
  codes_out[j] = codes_in[i];
}

__global__ __launch_bounds__(256) void k_iota(uint32_t *out, uint64_t n) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = (uint32_t)j;
}

}  // namespace

hipError_t launch_iota_u32(uint32_t *out, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, out, n);
  return hipGetLastError();
}

hipError_t launch_shoot_start(const ShootArgs &a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint64_t B = kSampleBlock;
  hipLaunchKernelGGL(k_shoot_start, dim3((uint32_t)((a.n + B - 1) / B)), dim3(B), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_shoot_bounce(const ShootArgs &a, uint64_t n_bound, hipStream_t stream) {
  if (n_bound == 0) return hipSuccess;
  const uint64_t B = kSampleBlock;
  hipLaunchKernelGGL(k_shoot_bounce, dim3((uint32_t)((n_bound + B - 1) / B)), dim3(B), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_photon_gather_records(const float *pos_in, const float *w_in, const uint32_t *codes_in,
                                        const uint32_t *order, uint64_t n, float *pos_out, float *w_out,
                                        uint32_t *codes_out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_photon_gather_records, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, pos_in, w_in,
                     codes_in, order, n, pos_out, w_out, codes_out);
  return hipGetLastError();
}

}  // namespace skw
