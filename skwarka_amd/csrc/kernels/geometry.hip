// Hit attributes in mesh space: the part of TriangleAccelerator::Intersect that follows the
// traversal (Source/RayLibs/Raytracer/Core/TriangleAccelerator.cpp:309-345) and the subset of
// TriangleMesh::ComputeDifferentialGeometry the hot path consumes (TriangleMesh.cpp:414-500):
// b1, b2, the point, geometric and shading normals, and m_cross / m_dot for GetNextMinT.
#include <hip/hip_runtime.h>

#include "geom_device.hpp"
#include "kernels.hpp"

namespace skw {
namespace {

using namespace dev;

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
  const DevMesh *mesh = a.meshes + p.mesh;
  const uint32_t *ti = a.mesh_tris + 3 * (mesh->tri_off + h.tri);
  const float *vb = a.mesh_verts + 3 * mesh->vert_off;
  D3 v[3];
  for (int k = 0; k < 3; ++k) v[k] = {(double)vb[3 * ti[k]], (double)vb[3 * ti[k] + 1], (double)vb[3 * ti[k] + 2]};
  D3 o, d;
  load_ray(a.rays, i, o, d);
  D3 mo = xpoint(p.w2m, o), md = xvec(p.w2m, d);
  double b1 = 0.0, b2 = 0.0, tm = 0.0;
  isect_point(v, mo, md, b1, b2, tm);
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
    D3 o, d;
    load_ray(a.rays, i, o, d);
    HitGeom hg = hit_geometry(a, o, d, h.prim, h.tri);
    g.p[0] = hg.p.x; g.p[1] = hg.p.y; g.p[2] = hg.p.z;
    g.ng[0] = hg.ng.x; g.ng[1] = hg.ng.y; g.ng[2] = hg.ng.z;
    g.ns[0] = hg.ns.x; g.ns[1] = hg.ns.y; g.ns[2] = hg.ns.z;
    g.cross[0] = hg.cross.x; g.cross[1] = hg.cross.y; g.cross[2] = hg.cross.z;
    g.dot = hg.dot;
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
