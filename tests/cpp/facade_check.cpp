// Drives the C++ classes of include/skwarka.hpp the way the reference's callers do (one ray, one
// lookup at a time), for the parity tests: reads a scene, rays, photons and gather points from a
// binary file, writes every result to another. Usage: facade_check IN OUT.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../include/skwarka.hpp"

namespace {

struct Reader {
  FILE *f;
  template <class T>
  T get() {
    T v;
    if (std::fread(&v, sizeof(T), 1, f) != 1) throw skw::Error("short input");
    return v;
  }
  template <class T>
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) throw skw::Error("short input");
    return v;
  }
};

struct Writer {
  FILE *f;
  template <class T>
  void put(const T &v) {
    std::fwrite(&v, sizeof(T), 1, f);
  }
};

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: facade_check IN OUT\n");
    return 2;
  }
  try {
    Reader in{std::fopen(argv[1], "rb")};
    if (!in.f) throw skw::Error("cannot open input");
    // meshes and primitives
    const uint32_t n_meshes = in.get<uint32_t>();
    std::vector<std::shared_ptr<const skw::TriangleMesh>> meshes;
    for (uint32_t m = 0; m < n_meshes; ++m) {
      const uint32_t nv = in.get<uint32_t>();
      std::vector<float> v = in.vec<float>(3 * (size_t)nv);
      const uint32_t nt = in.get<uint32_t>();
      std::vector<uint32_t> t = in.vec<uint32_t>(3 * (size_t)nt);
      const uint8_t invert = in.get<uint8_t>();
      std::vector<skw::Point3D_f> pts(nv);
      for (uint32_t i = 0; i < nv; ++i) pts[i] = skw::Point3D_f(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
      std::vector<skw::MeshTriangle> tris(nt);
      for (uint32_t i = 0; i < nt; ++i) tris[i] = skw::MeshTriangle{{t[3 * i], t[3 * i + 1], t[3 * i + 2]}};
      meshes.push_back(std::make_shared<skw::TriangleMesh>(std::move(pts), std::move(tris), invert != 0));
    }
    const uint32_t n_prims = in.get<uint32_t>();
    std::vector<std::shared_ptr<const skw::Primitive>> prims;
    for (uint32_t p = 0; p < n_prims; ++p) {
      const uint32_t mesh = in.get<uint32_t>();
      const uint8_t identity = in.get<uint8_t>();
      if (identity) {
        prims.push_back(std::make_shared<skw::Primitive>(meshes.at(mesh)));
      } else {
        std::vector<double> xf = in.vec<double>(32);
        prims.push_back(std::make_shared<skw::Primitive>(meshes.at(mesh), xf.data(), xf.data() + 16));
      }
    }
    const uint32_t n_rays = in.get<uint32_t>();
    std::vector<double> rays = in.vec<double>(8 * (size_t)n_rays);
    const uint32_t n_ph = in.get<uint32_t>();
    std::vector<float> pos = in.vec<float>(3 * (size_t)n_ph), w = in.vec<float>(3 * (size_t)n_ph);
    std::vector<uint16_t> dir = in.vec<uint16_t>(n_ph), nrm = in.vec<uint16_t>(n_ph);
    const uint32_t nq = in.get<uint32_t>();
    std::vector<double> qp = in.vec<double>(3 * (size_t)nq), qn = in.vec<double>(3 * (size_t)nq);
    const uint32_t k = in.get<uint32_t>();
    const double r = in.get<double>();
    std::fclose(in.f);

    Writer out{std::fopen(argv[2], "wb")};
    if (!out.f) throw skw::Error("cannot open output");
    skw::Scene scene(prims);
    skw::BBox3D_d b = scene.GetWorldBounds();
    for (int i = 0; i < 3; ++i) out.put(b.m_min[i]);
    for (int i = 0; i < 3; ++i) out.put(b.m_max[i]);
    for (uint32_t i = 0; i < n_rays; ++i) {
      const double *q = &rays[8 * (size_t)i];
      skw::Ray ray(skw::Point3D_d(q[0], q[1], q[2]), skw::Vector3D_d(q[3], q[4], q[5]), q[6], q[7]);
      skw::Intersection isect;
      double t = -1.0;
      const bool hit = scene.Intersect(skw::RayDifferential(ray), isect, &t);
      out.put<uint32_t>(hit ? (uint32_t)isect.m_primitive_index : 0xFFFFFFFFu);
      out.put<uint32_t>(hit ? (uint32_t)isect.m_triangle_index : 0xFFFFFFFFu);
      out.put(t);
      out.put(hit ? isect.m_b1 : 0.0);
      out.put(hit ? isect.m_b2 : 0.0);
      out.put<uint8_t>(scene.IntersectTest(ray) ? 1 : 0);
    }
    std::vector<skw::Photon> photons(n_ph);
    for (uint32_t i = 0; i < n_ph; ++i)
      photons[i] = skw::Photon(skw::Point3D_f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]), &w[3 * i], dir[i], nrm[i]);
    skw::KDTree<skw::Photon> tree(std::move(photons));
    out.put<uint64_t>(tree.GetNumberOfPoints());
    std::vector<skw::KDTree<skw::Photon>::NearestPoint> found(k);
    for (uint32_t i = 0; i < nq; ++i) {
      skw::Point3D_d p(qp[3 * i], qp[3 * i + 1], qp[3 * i + 2]);
      skw::Vector3D_d n(qn[3 * i], qn[3 * i + 1], qn[3 * i + 2]);
      skw::PhotonFilter filter(p, n, 0.87);
      // unfiltered and filtered k nearest, then the filtered nearest point
      for (int pass = 0; pass < 2; ++pass) {
        size_t c = pass == 0 ? tree.GetNearestPoints(p, k, found.data(), r)
                             : tree.GetNearestPoints(p, k, found.data(), filter, r);
        out.put<uint32_t>((uint32_t)c);
        for (uint32_t j = 0; j < k; ++j) {
          out.put<uint32_t>(j < c ? (uint32_t)tree.GetOriginalIndex(found[j].mp_point) : 0xFFFFFFFFu);
          out.put(j < c ? found[j].m_distance_sqr : 0.0);
        }
      }
      const skw::Photon *np = tree.GetNearestPoint(p, filter, r);
      out.put<uint32_t>(np ? (uint32_t)tree.GetOriginalIndex(np) : 0xFFFFFFFFu);
    }
    std::fclose(out.f);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "facade_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
