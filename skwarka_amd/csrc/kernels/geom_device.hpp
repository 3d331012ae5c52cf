// Device-side FP64 vector helpers and the mesh-space hit geometry shared by the hit-attribute
// and direct-lighting kernels. Operation order follows the reference's Vector3D (dot =
// (xx'+yy')+zz', cross, Normalize = multiply by 1/sqrt(len^2)) and Transform (row-major,
// (m0 x + m1 y) + m2 z + m3, w division when w != 1); every translation unit that includes this
// is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace skw {
namespace dev {

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 add3(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 mul3(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double len_sqr3(D3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ void normalize3(D3 &a) {
  double l2 = len_sqr3(a);
  if (l2 > 0.0) a = mul3(a, 1.0 / sqrt(l2));
}
__device__ __forceinline__ D3 xpoint(const double *m, D3 p) {
  double xp = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
  double yp = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
  double zp = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
  double w = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
  D3 r{xp, yp, zp};
  if (w != 1.0) r = mul3(r, 1.0 / w);
  return r;
}
__device__ __forceinline__ D3 xvec(const double *m, D3 v) {
  return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
          m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
// Transform::TransformNormal: transposed inverse, then normalised (Transform.h:281-290).
__device__ __forceinline__ D3 xnormal(const double *inv, D3 n) {
  D3 r{inv[0] * n.x + inv[4] * n.y + inv[8] * n.z, inv[1] * n.x + inv[5] * n.y + inv[9] * n.z,
       inv[2] * n.x + inv[6] * n.y + inv[10] * n.z};
  normalize3(r);
  return r;
}

// TriangleMesh::_ComputeIntersectionPoint (TriangleMesh.cpp:414-437).
__device__ __forceinline__ bool isect_point(const D3 v[3], D3 o, D3 d, double &b1, double &b2, double &t) {
  D3 e1 = sub3(v[1], v[0]), e2 = sub3(v[2], v[0]);
  D3 s1 = cross3(d, e2);
  double divisor = dot3(s1, e1);
  if (divisor == 0.0) return false;
  double inv = 1.0 / divisor;
  D3 dd = sub3(o, v[0]);
  b1 = dot3(dd, s1) * inv;
  D3 s2 = cross3(dd, e1);
  b2 = dot3(d, s2) * inv;
  t = dot3(e2, s2) * inv;
  return true;
}

struct HitGeom {
  D3 p, ng, ns, cross;
  double dot;
};

// World point, geometric and shading normals, and m_cross / m_dot of a closest hit
// (TriangleAccelerator.cpp:309-345, TriangleMesh.cpp:439-500).
__device__ __forceinline__ HitGeom hit_geometry(const GeomArgs &a, D3 o, D3 d, uint32_t prim, uint32_t tri) {
  const DevPrim &p = a.prims[prim];
  const DevMesh *mesh = a.meshes + p.mesh;
  const uint32_t *ti = a.mesh_tris + 3 * (mesh->tri_off + tri);
  const float *vb = a.mesh_verts + 3 * mesh->vert_off;
  D3 v[3];
  uint32_t idx[3];
  for (int k = 0; k < 3; ++k) {
    idx[k] = ti[k];
    v[k] = {(double)vb[3 * idx[k]], (double)vb[3 * idx[k] + 1], (double)vb[3 * idx[k] + 2]};
  }
  D3 mo = xpoint(p.w2m, o), md = xvec(p.w2m, d);
  double b1 = 0.0, b2 = 0.0, tm = 0.0;
  isect_point(v, mo, md, b1, b2, tm);
  D3 pm = add3(mo, mul3(md, tm));
  HitGeom g;
  g.p = xpoint(p.m2w, pm);
  const bool inv_n = mesh->invert_normals != 0;
  D3 gn = inv_n ? cross3(sub3(v[2], v[0]), sub3(v[1], v[0])) : cross3(sub3(v[1], v[0]), sub3(v[2], v[0]));
  normalize3(gn);
  D3 sn = gn;
  if (mesh->has_normals) {
    double b0 = 1.0 - b1 - b2;
    const float *nb = a.mesh_normals + 3 * mesh->vert_off;
    D3 vn[3];
    for (int k = 0; k < 3; ++k) {
      vn[k] = {(double)nb[3 * idx[k]], (double)nb[3 * idx[k] + 1], (double)nb[3 * idx[k] + 2]};
      if (inv_n) vn[k] = mul3(vn[k], -1.0);
    }
    sn = add3(add3(mul3(vn[0], b0), mul3(vn[1], b1)), mul3(vn[2], b2));
    normalize3(sn);
  }
  g.ng = xnormal(p.w2m, gn);
  g.ns = xnormal(p.w2m, sn);
  g.cross = cross3(sub3(v[1], v[0]), sub3(v[2], v[0]));
  D3 pback = xpoint(p.w2m, g.p);
  g.dot = dot3(cross3(sub3(v[0], pback), sub3(v[1], v[0])), sub3(v[2], v[0]));
  return g;
}

// CoreUtils::GetNextMinT (CoreUtils.h:139-149).
__device__ __forceinline__ double next_min_t(const DevPrim &p, const HitGeom &g, D3 dir) {
  D3 dm = xvec(p.w2m, dir);
  double divisor = dot3(dm, g.cross);
  if (divisor == 0.0) return 0.0;
  double v = (g.dot * (1.0 / divisor)) + 1e-14;
  return (v > 0.0) ? v : 0.0;
}

}  // namespace dev
}  // namespace skw
