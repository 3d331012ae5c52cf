// skwarka.hpp — C++ classes with the reference's names and signatures over the skwarka_rt C ABI,
// so a caller of Skwarka's scene / accelerator / photon-tree interfaces can switch by include:
//
//   TriangleAccelerator(std::vector<...Primitive>), Intersect, IntersectTest, GetWorldBounds
//       Source/RayLibs/Raytracer/Core/TriangleAccelerator.h:41-62
//   Scene::GetWorldBounds / Intersect / IntersectTest
//       Source/RayLibs/Raytracer/Core/Scene.h:68-85
//   KDTree<Photon>(points), GetNumberOfPoints, GetAllPoints, GetNearestPoint, GetNearestPoints,
//   NearestPoint {mp_point, m_distance_sqr}
//       Source/RayLibs/Raytracer/Core/KDTree.h:40-119, 163-177
//   Photon {m_point, m_weight, m_incident_direction, m_normal}, PhotonFilter(point, normal, cos)
//       Source/RayLibs/Raytracer/LTEIntegrators/PhotonLTEIntegrator/PhotonInternalTypes.h:35-59, 113-142
//
// Scalar calls run as a batch of one on the device (a few tens of microseconds each); callers with
// many rays use the batched overloads, which take device arrays and a HIP stream. Objects are
// immutable after construction and may be used from several threads. Errors throw skw::Error
// (the reference asserts in debug builds and has no error channel).
//
// Differences, all forced by the boundary: primitives are geometry only (mesh + transform);
// Intersection carries the hit primitive's index and the triangle index instead of a Primitive
// pointer; KDTree is provided for Photon; PhotonFilter must be centred at the query point with
// cos threshold 0.87 (the photon integrator's only use).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "skwarka_rt.h"

namespace skw {

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string &what) : std::runtime_error(what) {}
};

inline void check(skw_status s, const char *what) {
  if (s != SKW_OK) throw Error(std::string(what) + ": " + skw_last_error());
}
inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- value types (reference names)
template <class T>
struct Point3D {
  T m_x = 0, m_y = 0, m_z = 0;
  Point3D() = default;
  Point3D(T x, T y, T z) : m_x(x), m_y(y), m_z(z) {}
  T operator[](unsigned i) const { return i == 0 ? m_x : (i == 1 ? m_y : m_z); }
};
template <class T>
struct Vector3D {
  T m_x = 0, m_y = 0, m_z = 0;
  Vector3D() = default;
  Vector3D(T x, T y, T z) : m_x(x), m_y(y), m_z(z) {}
  T operator[](unsigned i) const { return i == 0 ? m_x : (i == 1 ? m_y : m_z); }
};
using Point3D_d = Point3D<double>;
using Point3D_f = Point3D<float>;
using Vector3D_d = Vector3D<double>;

struct BBox3D_d {
  Point3D_d m_min, m_max;
};

// Ray.h:31-60: origin, normalised direction, [min_t, max_t].
struct Ray {
  Point3D_d m_origin;
  Vector3D_d m_direction;
  double m_min_t = 0.0, m_max_t = std::numeric_limits<double>::infinity();
  Ray() = default;
  Ray(const Point3D_d &o, const Vector3D_d &d, double min_t = 0.0,
      double max_t = std::numeric_limits<double>::infinity())
      : m_origin(o), m_direction(d), m_min_t(min_t), m_max_t(max_t) {}
};

// RayDifferential.h: the base ray plus (unused here) differentials.
struct RayDifferential {
  Ray m_base_ray;
  bool m_has_differentials = false;
  RayDifferential() = default;
  explicit RayDifferential(const Ray &r) : m_base_ray(r) {}
};

struct DifferentialGeometry {
  Point3D_d m_point;
  Vector3D_d m_geometric_normal;
  struct {
    Vector3D_d m_normal;
  } m_shading;
};

struct Intersection {
  DifferentialGeometry m_dg;
  size_t m_primitive_index = 0;  // index into the scene's primitive list
  size_t m_triangle_index = 0;   // triangle inside the primitive's mesh
  double m_b1 = 0.0, m_b2 = 0.0;
  double m_cross[3] = {0, 0, 0}, m_dot = 0.0;  // GetNextMinT inputs (CoreUtils.h:139-149)
};

// ---------------------------------------------------------------- geometry
struct MeshTriangle {
  uint32_t m_vertices[3];
};

// TriangleMesh: float vertices, index triples, optional shading normals, invert flag.
class TriangleMesh {
 public:
  TriangleMesh(std::vector<Point3D_f> vertices, std::vector<MeshTriangle> triangles, bool invert_normals = false,
               std::vector<Vector3D<float>> shading_normals = {})
      : m_vertices(std::move(vertices)),
        m_triangles(std::move(triangles)),
        m_normals(std::move(shading_normals)),
        m_invert_normals(invert_normals) {}
  const std::vector<Point3D_f> &GetVertices() const { return m_vertices; }
  const std::vector<MeshTriangle> &GetTriangles() const { return m_triangles; }
  const std::vector<Vector3D<float>> &GetShadingNormals() const { return m_normals; }
  bool GetInvertNormals() const { return m_invert_normals; }

 private:
  std::vector<Point3D_f> m_vertices;
  std::vector<MeshTriangle> m_triangles;
  std::vector<Vector3D<float>> m_normals;
  bool m_invert_normals;
};

// Primitive: a mesh placed in the world by mesh_to_world (row-major 4x4) and its inverse.
class Primitive {
 public:
  explicit Primitive(std::shared_ptr<const TriangleMesh> mesh) : mp_mesh(std::move(mesh)) {
    for (int i = 0; i < 16; ++i) m_mesh_to_world[i] = m_world_to_mesh[i] = (i % 5 == 0) ? 1.0 : 0.0;
    m_identity = true;
  }
  Primitive(std::shared_ptr<const TriangleMesh> mesh, const double mesh_to_world[16], const double world_to_mesh[16])
      : mp_mesh(std::move(mesh)) {
    std::memcpy(m_mesh_to_world, mesh_to_world, sizeof(m_mesh_to_world));
    std::memcpy(m_world_to_mesh, world_to_mesh, sizeof(m_world_to_mesh));
    m_identity = false;
  }
  const std::shared_ptr<const TriangleMesh> &GetTriangleMesh() const { return mp_mesh; }
  const double *GetMeshToWorld() const { return m_mesh_to_world; }
  const double *GetWorldToMesh() const { return m_world_to_mesh; }
  bool IsIdentity() const { return m_identity; }

 private:
  std::shared_ptr<const TriangleMesh> mp_mesh;
  double m_mesh_to_world[16], m_world_to_mesh[16];
  bool m_identity;
};

namespace detail {

// Device buffer owned by one scalar call.
class DeviceBuffer {
 public:
  explicit DeviceBuffer(size_t bytes) { check_hip(hipMalloc(&p_, bytes ? bytes : 1), "hipMalloc"); }
  ~DeviceBuffer() { (void)hipFree(p_); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  template <class T>
  T *as() const {
    return static_cast<T *>(p_);
  }
  void put(const void *src, size_t bytes) { check_hip(hipMemcpy(p_, src, bytes, hipMemcpyHostToDevice), "upload"); }
  void get(void *dst, size_t bytes) const {
    check_hip(hipMemcpy(dst, p_, bytes, hipMemcpyDeviceToHost), "download");
  }

 private:
  void *p_ = nullptr;
};

inline skw_ray to_abi(const Ray &r) {
  skw_ray o;
  o.o[0] = r.m_origin.m_x;
  o.o[1] = r.m_origin.m_y;
  o.o[2] = r.m_origin.m_z;
  o.d[0] = r.m_direction.m_x;
  o.d[1] = r.m_direction.m_y;
  o.d[2] = r.m_direction.m_z;
  o.tmin = r.m_min_t;
  o.tmax = r.m_max_t;
  return o;
}

struct SceneDeleter {
  void operator()(skw_scene *s) const { skw_scene_destroy(s); }
};

}  // namespace detail

// ---------------------------------------------------------------- TriangleAccelerator
class TriangleAccelerator {
 public:
  // TriangleAccelerator.h:41: builds the ternary tree over the primitives (shared meshes are
  // instanced) and uploads it to the current HIP device.
  explicit TriangleAccelerator(std::vector<std::shared_ptr<const Primitive>> i_primitives) {
    std::vector<const TriangleMesh *> meshes;
    std::vector<uint32_t> prim_mesh;
    for (const auto &p : i_primitives) {
      const TriangleMesh *m = p->GetTriangleMesh().get();
      uint32_t id = 0;
      while (id < meshes.size() && meshes[id] != m) ++id;
      if (id == meshes.size()) meshes.push_back(m);
      prim_mesh.push_back(id);
    }
    std::vector<uint64_t> voff{0}, toff{0};
    std::vector<float> verts, normals;
    std::vector<uint32_t> tris;
    std::vector<uint8_t> flags;
    bool any_normals = false;
    for (const TriangleMesh *m : meshes) any_normals |= !m->GetShadingNormals().empty();
    for (const TriangleMesh *m : meshes) {
      for (const auto &v : m->GetVertices()) {
        verts.push_back(v.m_x);
        verts.push_back(v.m_y);
        verts.push_back(v.m_z);
      }
      const auto &n = m->GetShadingNormals();
      if (any_normals)
        for (size_t i = 0; i < m->GetVertices().size(); ++i) {
          normals.push_back(n.empty() ? 0.0f : n[i].m_x);
          normals.push_back(n.empty() ? 0.0f : n[i].m_y);
          normals.push_back(n.empty() ? 0.0f : n[i].m_z);
        }
      for (const auto &t : m->GetTriangles())
        for (int k = 0; k < 3; ++k) tris.push_back(t.m_vertices[k]);
      voff.push_back(verts.size() / 3);
      toff.push_back(tris.size() / 3);
      flags.push_back((uint8_t)((m->GetInvertNormals() ? SKW_MESH_INVERT_NORMALS : 0u) |
                                (n.empty() ? 0u : SKW_MESH_SHADING_NORMALS)));
    }
    std::vector<double> xf;
    bool all_identity = true;
    for (const auto &p : i_primitives) all_identity &= p->IsIdentity();
    if (!all_identity)
      for (const auto &p : i_primitives) {
        xf.insert(xf.end(), p->GetMeshToWorld(), p->GetMeshToWorld() + 16);
        xf.insert(xf.end(), p->GetWorldToMesh(), p->GetWorldToMesh() + 16);
      }
    skw_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_meshes = (uint32_t)meshes.size();
    d.mesh_vert_offset = voff.data();
    d.verts = verts.data();
    d.normals = any_normals ? normals.data() : nullptr;
    d.mesh_tri_offset = toff.data();
    d.tris = tris.data();
    d.mesh_flags = flags.data();
    d.n_prims = (uint32_t)i_primitives.size();
    d.prim_mesh = prim_mesh.data();
    d.prim_xform = all_identity ? nullptr : xf.data();
    skw_scene *s = nullptr;
    check(skw_scene_create(&d, 0, &s), "skw_scene_create");
    mp_scene.reset(s, detail::SceneDeleter());
  }

  // TriangleAccelerator.h:50 (batch of one): closest hit with its differential geometry.
  bool Intersect(const RayDifferential &i_ray, Intersection &o_intersection, double *o_t = NULL) const {
    skw_ray r = detail::to_abi(i_ray.m_base_ray);
    detail::DeviceBuffer b(sizeof(skw_ray) + sizeof(skw_hit) + sizeof(skw_hit_geom));
    skw_ray *d_ray = b.as<skw_ray>();
    skw_hit *d_hit = reinterpret_cast<skw_hit *>(d_ray + 1);
    skw_hit_geom *d_geom = reinterpret_cast<skw_hit_geom *>(d_hit + 1);
    b.put(&r, sizeof(r));
    check(skw_closest_hit(mp_scene.get(), d_ray, 1, d_hit, nullptr), "skw_closest_hit");
    check(skw_hit_geometry(mp_scene.get(), d_ray, d_hit, 1, d_geom, nullptr), "skw_hit_geometry");
    char host[sizeof(skw_ray) + sizeof(skw_hit) + sizeof(skw_hit_geom)];
    b.get(host, sizeof(host));
    skw_hit h;
    skw_hit_geom g;
    std::memcpy(&h, host + sizeof(skw_ray), sizeof(h));
    std::memcpy(&g, host + sizeof(skw_ray) + sizeof(skw_hit), sizeof(g));
    if (o_t) *o_t = h.t;
    if (h.prim == 0xFFFFFFFFu) return false;
    o_intersection.m_primitive_index = h.prim;
    o_intersection.m_triangle_index = h.tri;
    o_intersection.m_b1 = h.b1;
    o_intersection.m_b2 = h.b2;
    o_intersection.m_dg.m_point = Point3D_d(g.p[0], g.p[1], g.p[2]);
    o_intersection.m_dg.m_geometric_normal = Vector3D_d(g.ng[0], g.ng[1], g.ng[2]);
    o_intersection.m_dg.m_shading.m_normal = Vector3D_d(g.ns[0], g.ns[1], g.ns[2]);
    for (int i = 0; i < 3; ++i) o_intersection.m_cross[i] = g.cross[i];
    o_intersection.m_dot = g.dot;
    return true;
  }

  // TriangleAccelerator.h:57 (batch of one): any hit inside [min_t, max_t].
  bool IntersectTest(const Ray &i_ray) const {
    skw_ray r = detail::to_abi(i_ray);
    detail::DeviceBuffer b(sizeof(skw_ray) + 8);
    b.put(&r, sizeof(r));
    uint8_t *d_occ = reinterpret_cast<uint8_t *>(b.as<skw_ray>() + 1);
    check(skw_any_hit(mp_scene.get(), b.as<skw_ray>(), 1, d_occ, nullptr), "skw_any_hit");
    uint8_t host[sizeof(skw_ray) + 8];
    b.get(host, sizeof(host));
    return host[sizeof(skw_ray)] != 0;
  }

  // TriangleAccelerator.h:62
  BBox3D_d GetWorldBounds() const {
    double mm[6];
    check(skw_scene_bounds(mp_scene.get(), mm), "skw_scene_bounds");
    BBox3D_d b;
    b.m_min = Point3D_d(mm[0], mm[1], mm[2]);
    b.m_max = Point3D_d(mm[3], mm[4], mm[5]);
    return b;
  }

  // Batched forms (device arrays, stream-ordered).
  void Intersect(const skw_ray *d_rays, size_t n, skw_hit *d_hits, hipStream_t stream = nullptr) const {
    check(skw_closest_hit(mp_scene.get(), d_rays, n, d_hits, stream), "skw_closest_hit");
  }
  void IntersectTest(const skw_ray *d_rays, size_t n, uint8_t *d_occluded, hipStream_t stream = nullptr) const {
    check(skw_any_hit(mp_scene.get(), d_rays, n, d_occluded, stream), "skw_any_hit");
  }
  void HitGeometry(const skw_ray *d_rays, const skw_hit *d_hits, size_t n, skw_hit_geom *d_geom,
                   hipStream_t stream = nullptr) const {
    check(skw_hit_geometry(mp_scene.get(), d_rays, d_hits, n, d_geom, stream), "skw_hit_geometry");
  }

  const skw_scene *GetHandle() const { return mp_scene.get(); }

 private:
  std::shared_ptr<skw_scene> mp_scene;
};

// ---------------------------------------------------------------- Scene
class Scene {
 public:
  explicit Scene(std::vector<std::shared_ptr<const Primitive>> i_primitives)
      : m_primitives(i_primitives), m_triangle_accelerator(std::move(i_primitives)) {}

  const std::vector<std::shared_ptr<const Primitive>> &GetPrimitives() const { return m_primitives; }
  // Scene.h:68
  BBox3D_d GetWorldBounds() const { return m_triangle_accelerator.GetWorldBounds(); }
  // Scene.h:77: o_t is set to infinity on a miss.
  bool Intersect(const RayDifferential &i_ray, Intersection &o_intersection, double *o_t = NULL) const {
    return m_triangle_accelerator.Intersect(i_ray, o_intersection, o_t);
  }
  // Scene.h:85
  bool IntersectTest(const Ray &i_ray) const { return m_triangle_accelerator.IntersectTest(i_ray); }
  const TriangleAccelerator &GetAccelerator() const { return m_triangle_accelerator; }

 private:
  std::vector<std::shared_ptr<const Primitive>> m_primitives;
  TriangleAccelerator m_triangle_accelerator;
};

// ---------------------------------------------------------------- photons
// PhotonInternalTypes.h:35-59: 28 bytes, direction and normal as CompressedDirection ids.
struct Photon {
  Point3D_f m_point;
  float m_weight[3] = {0, 0, 0};
  uint16_t m_incident_direction = 0, m_normal = 0;
  Photon() = default;
  Photon(const Point3D_f &p, const float w[3], uint16_t dir, uint16_t normal)
      : m_point(p), m_incident_direction(dir), m_normal(normal) {
    for (int i = 0; i < 3; ++i) m_weight[i] = w[i];
  }
  float operator[](unsigned char i) const { return m_point[i]; }
};

// PhotonInternalTypes.h:113-142. The device kernels apply it with the query point as surface point
// and cos threshold 0.87.
class PhotonFilter {
 public:
  PhotonFilter(const Point3D_d &i_surface_point, const Vector3D_d &i_surface_normal, double i_cos_threshold)
      : m_surface_point(i_surface_point), m_surface_normal(i_surface_normal), m_cos_threshold(i_cos_threshold) {}
  const Point3D_d &GetSurfacePoint() const { return m_surface_point; }
  const Vector3D_d &GetSurfaceNormal() const { return m_surface_normal; }
  double GetCosThreshold() const { return m_cos_threshold; }

 private:
  Point3D_d m_surface_point;
  Vector3D_d m_surface_normal;
  double m_cos_threshold;
};

template <class TPoint3D>
class KDTree;

// KDTree<Photon> (KDTree.h:40-119): the tree stores the points in tree order, as the reference's
// in-place build leaves them, so GetAllPoints() and NearestPoint::mp_point refer to that order.
template <>
class KDTree<Photon> {
 public:
  struct NearestPoint {
    const Photon *mp_point = nullptr;
    double m_distance_sqr = 0.0;
    bool operator<(const NearestPoint &o) const {
      return m_distance_sqr == o.m_distance_sqr ? (mp_point < o.mp_point) : (m_distance_sqr < o.m_distance_sqr);
    }
  };

  explicit KDTree(const std::vector<Photon> &i_points) { build(i_points); }
  explicit KDTree(std::vector<Photon> &&i_points) { build(i_points); }

  size_t GetNumberOfPoints() const { return m_points.size(); }
  const std::vector<Photon> &GetAllPoints() const { return m_points; }

  const Photon *GetNearestPoint(const Point3D_d &i_point,
                                double i_max_distance = std::numeric_limits<double>::infinity()) const {
    return nearest(i_point, nullptr, i_max_distance);
  }
  const Photon *GetNearestPoint(const Point3D_d &i_point, const PhotonFilter &i_filter,
                                double i_max_distance = std::numeric_limits<double>::infinity()) const {
    return nearest(i_point, &i_filter, i_max_distance);
  }
  size_t GetNearestPoints(const Point3D_d &i_point, size_t i_points_to_lookup, NearestPoint *op_nearest_points,
                          double i_max_distance = std::numeric_limits<double>::infinity()) const {
    return knn(i_point, i_points_to_lookup, op_nearest_points, nullptr, i_max_distance);
  }
  size_t GetNearestPoints(const Point3D_d &i_point, size_t i_points_to_lookup, NearestPoint *op_nearest_points,
                          const PhotonFilter &i_filter,
                          double i_max_distance = std::numeric_limits<double>::infinity()) const {
    return knn(i_point, i_points_to_lookup, op_nearest_points, &i_filter, i_max_distance);
  }

  const skw_photon_map *GetHandle() const { return mp_map.get(); }
  // Position, in the constructor's input, of a point of GetAllPoints().
  size_t GetOriginalIndex(const Photon *p) const { return m_perm[(size_t)(p - m_points.data())]; }

 private:
  struct MapDeleter {
    void operator()(skw_photon_map *m) const { skw_photon_map_destroy(m); }
  };

  void build(const std::vector<Photon> &pts) {
    const size_t n = pts.size();
    std::vector<float> pos(3 * n), w(3 * n);
    std::vector<uint16_t> dir(n), nrm(n);
    for (size_t i = 0; i < n; ++i) {
      pos[3 * i] = pts[i].m_point.m_x;
      pos[3 * i + 1] = pts[i].m_point.m_y;
      pos[3 * i + 2] = pts[i].m_point.m_z;
      for (int c = 0; c < 3; ++c) w[3 * i + c] = pts[i].m_weight[c];
      dir[i] = pts[i].m_incident_direction;
      nrm[i] = pts[i].m_normal;
    }
    skw_photon_map *m = nullptr;
    check(skw_photon_map_create(n, pos.data(), w.data(), dir.data(), nrm.data(), 0, &m), "skw_photon_map_create");
    mp_map.reset(m, MapDeleter());
    m_perm.assign(n, 0);
    if (n > 0) check(skw_photon_map_export(m, m_perm.data(), nullptr, nullptr), "skw_photon_map_export");
    m_points.resize(n);
    m_tree_pos.resize(n);
    for (size_t t = 0; t < n; ++t) {
      m_points[t] = pts[m_perm[t]];
      m_tree_pos[m_perm[t]] = (uint32_t)t;
    }
  }

  static int filter_mode(const Point3D_d &q, const PhotonFilter *f) {
    if (!f) return 0;
    const Point3D_d &s = f->GetSurfacePoint();
    if (s.m_x != q.m_x || s.m_y != q.m_y || s.m_z != q.m_z || f->GetCosThreshold() != 0.87)
      throw Error("PhotonFilter: the device filter is centred at the query point with cos threshold 0.87");
    return 1;
  }

  size_t knn(const Point3D_d &q, size_t k, NearestPoint *out, const PhotonFilter *f, double r) const {
    if (k == 0 || m_points.empty()) return 0;
    const int mode = filter_mode(q, f);
    const double qn[6] = {q.m_x, q.m_y, q.m_z, f ? f->GetSurfaceNormal().m_x : 0.0,
                          f ? f->GetSurfaceNormal().m_y : 0.0, f ? f->GetSurfaceNormal().m_z : 0.0};
    detail::DeviceBuffer b(6 * sizeof(double) + k * (sizeof(uint32_t) + sizeof(double)) + 8);
    b.put(qn, sizeof(qn));
    double *d_q = b.as<double>();
    double *d_d2 = d_q + 6;
    uint32_t *d_ids = reinterpret_cast<uint32_t *>(d_d2 + k);
    uint32_t *d_count = d_ids + k;
    check(skw_photon_knn(mp_map.get(), d_q, mode ? d_q + 3 : nullptr, 1, (uint32_t)k, r, mode, d_ids, d_d2, d_count,
                         nullptr),
          "skw_photon_knn");
    std::vector<char> host(6 * sizeof(double) + k * (sizeof(uint32_t) + sizeof(double)) + 4);
    b.get(host.data(), host.size());
    const double *d2 = reinterpret_cast<const double *>(host.data() + 6 * sizeof(double));
    const uint32_t *ids = reinterpret_cast<const uint32_t *>(d2 + k);
    const uint32_t count = ids[k];
    for (uint32_t j = 0; j < count; ++j) {
      out[j].mp_point = &m_points[m_tree_pos[ids[j]]];
      out[j].m_distance_sqr = d2[j];
    }
    return count;
  }

  const Photon *nearest(const Point3D_d &q, const PhotonFilter *f, double r) const {
    if (m_points.empty()) return nullptr;
    const int mode = filter_mode(q, f);
    const double qn[6] = {q.m_x, q.m_y, q.m_z, f ? f->GetSurfaceNormal().m_x : 0.0,
                          f ? f->GetSurfaceNormal().m_y : 0.0, f ? f->GetSurfaceNormal().m_z : 0.0};
    detail::DeviceBuffer b(6 * sizeof(double) + 8);
    b.put(qn, sizeof(qn));
    uint32_t *d_id = reinterpret_cast<uint32_t *>(b.as<double>() + 6);
    check(skw_photon_nearest(mp_map.get(), b.as<double>(), mode ? b.as<double>() + 3 : nullptr, 1, r, mode, d_id,
                             nullptr),
          "skw_photon_nearest");
    char host[6 * sizeof(double) + 8];
    b.get(host, sizeof(host));
    uint32_t id;
    std::memcpy(&id, host + 6 * sizeof(double), sizeof(id));
    return id == 0xFFFFFFFFu ? nullptr : &m_points[m_tree_pos[id]];
  }

  std::shared_ptr<skw_photon_map> mp_map;
  std::vector<Photon> m_points;     // tree order
  std::vector<uint32_t> m_tree_pos;  // caller index -> tree position
  std::vector<uint32_t> m_perm;      // tree position -> caller index
};

}  // namespace skw
