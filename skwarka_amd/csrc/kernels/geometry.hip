// Hit attributes in mesh space: the part of TriangleAccelerator::Intersect that follows the
// traversal (Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:309-345) and the subset of
// TriangleMesh::ComputeDifferentialGeometry the hot path consumes (TriangleMesh.cpp:414-500):
// b1, b2, the point, geometric and shading normals, and m_cross / m_dot for GetNextMinT.
// Compiled with -ffp-contract=off; operation order follows Vector3D (dot = (xx'+yy')+zz').
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace skw {
namespace {

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
__device__ __forceinline__ void normalize3(D3 &a) {
  double l2 = a.x * a.x + a.y * a.y + a.z * a.z;
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

struct MeshTri {
  D3 v[3];
  uint32_t idx[3];
  const DevMesh *mesh;
};

__device__ __forceinline__ MeshTri load_mesh_tri(const GeomArgs &a, const DevPrim &p, uint32_t tri) {
  MeshTri t;
  t.mesh = a.meshes + p.mesh;
  const uint32_t *ti = a.mesh_tris + 3 * (t.mesh->tri_off + tri);
  const float *vb = a.mesh_verts + 3 * t.mesh->vert_off;
  for (int k = 0; k < 3; ++k) {
    t.idx[k] = ti[k];
    t.v[k] = {(double)vb[3 * ti[k]], (double)vb[3 * ti[k] + 1], (double)vb[3 * ti[k] + 2]};
  }
  return t;
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

__device__ __forceinline__ void load_ray(const skw_ray *rays, uint64_t i, D3 &o, D3 &d) {
  const skw_ray &r = rays[i];
  o = {r.o[0], r.o[1], r.o[2]};
  d = {r.d[0], r.d[1], r.d[2]};
}

__global__ __launch_bounds__(256) void k_fix_bary(GeomArgs a) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  skw_hit h = a.hits[i];
  if (h.prim == kMiss) return;
  const DevPrim &p = a.prims[h.prim];
  if (p.identity) return;  // traversal barycentrics already are the mesh-space ones
  MeshTri t = load_mesh_tri(a, p, h.tri);
  D3 o, d;
  load_ray(a.rays, i, o, d);
  D3 mo = xpoint(p.w2m, o), md = xvec(p.w2m, d);
  double b1 = 0.0, b2 = 0.0, tm = 0.0;
  isect_point(t.v, mo, md, b1, b2, tm);
  a.hits[i].b1 = b1;
  a.hits[i].b2 = b2;
}

__global__ __launch_bounds__(256) void k_hit_geometry(GeomArgs a) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  skw_hit h = a.hits[i];
  skw_hit_geom g;
  for (int c = 0; c < 3; ++c) g.p[c] = g.ng[c] = g.ns[c] = g.cross[c] = 0.0;
  g.dot = 0.0;
  if (h.prim != kMiss) {
    const DevPrim &p = a.prims[h.prim];
    MeshTri t = load_mesh_tri(a, p, h.tri);
    D3 o, d;
    load_ray(a.rays, i, o, d);
    D3 mo = xpoint(p.w2m, o), md = xvec(p.w2m, d);
    double b1 = 0.0, b2 = 0.0, tm = 0.0;
    isect_point(t.v, mo, md, b1, b2, tm);
    D3 pm = add3(mo, mul3(md, tm));
    D3 pw = xpoint(p.m2w, pm);
    bool inv_n = t.mesh->invert_normals != 0;
    D3 gn = inv_n ? cross3(sub3(t.v[2], t.v[0]), sub3(t.v[1], t.v[0])) : cross3(sub3(t.v[1], t.v[0]), sub3(t.v[2], t.v[0]));
    normalize3(gn);
    D3 sn = gn;
    if (t.mesh->has_normals) {
      double b0 = 1.0 - b1 - b2;
      const float *nb = a.mesh_normals + 3 * t.mesh->vert_off;
      D3 vn[3];
      for (int k = 0; k < 3; ++k) {
        vn[k] = {(double)nb[3 * t.idx[k]], (double)nb[3 * t.idx[k] + 1], (double)nb[3 * t.idx[k] + 2]};
        if (inv_n) vn[k] = mul3(vn[k], -1.0);
      }
      sn = add3(add3(mul3(vn[0], b0), mul3(vn[1], b1)), mul3(vn[2], b2));
      normalize3(sn);
    }
    D3 gw = xnormal(p.w2m, gn), sw = xnormal(p.w2m, sn);
    D3 cr = cross3(sub3(t.v[1], t.v[0]), sub3(t.v[2], t.v[0]));
    D3 pback = xpoint(p.w2m, pw);
    double dt = dot3(cross3(sub3(t.v[0], pback), sub3(t.v[1], t.v[0])), sub3(t.v[2], t.v[0]));
    g.p[0] = pw.x; g.p[1] = pw.y; g.p[2] = pw.z;
    g.ng[0] = gw.x; g.ng[1] = gw.y; g.ng[2] = gw.z;
    g.ns[0] = sw.x; g.ns[1] = sw.y; g.ns[2] = sw.z;
    g.cross[0] = cr.x; g.cross[1] = cr.y; g.cross[2] = cr.z;
    g.dot = dt;
  }
  a.geom[i] = g;
}

}  // namespace

hipError_t launch_fix_bary(const GeomArgs &args, hipStream_t stream) {
  if (args.n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)((args.n + 255) / 256);
  hipLaunchKernelGGL(k_fix_bary, dim3(grid), dim3(256), 0, stream, args);
  return hipGetLastError();
}

hipError_t launch_hit_geometry(const GeomArgs &args, hipStream_t stream) {
  if (args.n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)((args.n + 255) / 256);
  hipLaunchKernelGGL(k_hit_geometry, dim3(grid), dim3(256), 0, stream, args);
  return hipGetLastError();
}

}  // namespace skw
