// ORACLE — TEST INFRASTRUCTURE ONLY (see orc_math.hpp). Never linked into the product library.
//
// CPU restatement of Skwarka's point kd-tree and the photon-map density estimators:
//   KDTree<T>::_Build (longest-axis median split via std::nth_element)  Source/RayLibs/Raytracer/Core/KDTree.h:427-473
//   Node packing (float split, 29-bit right child, has-left bit, axis)  KDTree.h:187-246
//   _Lookup (recursive, strict d^2 < max, near side first)              KDTree.h:535-576
//   NearestPointsProc (fill, make_heap, pop/replace/push)               KDTree.h:331-387
//   NearestPoint::operator< ((d^2, address) order)                      KDTree.h:174-177
//   NearestPointProc (filtered nearest)                                 KDTree.h:293-323
//   Photon (28 B: point, weight, 2 compressed directions)               Raytracer/LTEIntegrators/PhotonLTEIntegrator/PhotonInternalTypes.h:35-59
//   PhotonFilter (normal within 0.87 cosine, disc-plane test)           PhotonInternalTypes.h:113-142
//   MAX_NORMAL_DEVIATION_COS = 0.87                                      PhotonLTEIntegrator.cpp:27
//   _LookupPhotonIrradiance (uniform disc, n/(n-0.5) correction)        PhotonLTEIntegrator.cpp:127-169
//   _LookupCausticRadiance + _PhotonKernel (Simpson kernel)             PhotonLTEIntegrator.cpp:579-636
//   CompressedDirection codec and decode table                          Source/RayLibs/Math/CompressedDirection.h:92-124, .cpp:20-44
// The heap operations are libstdc++'s std::make_heap / pop_heap / push_heap: the final array
// order (and so the summation order of the estimators) is that library's.
#include <algorithm>
#include <cmath>
#include <functional>
#include <random>
#include <vector>

#include "../include/orc_api.h"
#include "orc_math.hpp"

namespace orc {

static double g_cdir[1 << 16][3];
static bool g_cdir_ready = false;

static void cdir_init() {
  if (g_cdir_ready) return;
  for (int i = 0; i < (1 << 16); ++i) {
    unsigned char x = (unsigned char)((i & 0x1FC0) >> 6);
    unsigned char y = (unsigned char)(i & 0x3F);
    if (x + y >= 127) {
      x = (unsigned char)(127 - x);
      y = (unsigned char)(127 - y);
    }
    V3 d{x + 0.5, y + 0.5, (double)(126 - x - y)};
    normalize(d);
    if (i & (1 << 13)) d.x = -d.x;
    if (i & (1 << 14)) d.y = -d.y;
    if (i & (1 << 15)) d.z = -d.z;
    g_cdir[i][0] = d.x;
    g_cdir[i][1] = d.y;
    g_cdir[i][2] = d.z;
  }
  g_cdir_ready = true;
}
static inline V3 cdir(uint16_t c) { return V3{g_cdir[c][0], g_cdir[c][1], g_cdir[c][2]}; }

static uint16_t cdir_encode(V3 v) {
  unsigned int data = 0;
  if (v.x < 0) { data |= 1u << 13; v.x = -v.x; }
  if (v.y < 0) { data |= 1u << 14; v.y = -v.y; }
  if (v.z < 0) { data |= 1u << 15; v.z = -v.z; }
  double scale = 126.999 / (v.x + v.y + v.z);
  unsigned char x = (unsigned char)(v.x * scale);
  unsigned char y = (unsigned char)(v.y * scale);
  if (y >= 64) {
    x = (unsigned char)(127 - x);
    y = (unsigned char)(127 - y);
  }
  data |= ((unsigned int)x << 6) | y;
  return (uint16_t)data;
}

struct NodeK {
  float split = 0.0f;
  uint32_t word = 0;  // bits 0-1 axis (3 = leaf), bit 2 has-left, bits 3..31 right child
  void set_right(uint32_t r) { word = (r << 3) + (word & 7); }
  uint32_t right() const { return word >> 3; }
  void set_has_left(bool b) { word = (word & ~4u) + (b ? 4u : 0u); }
  bool has_left() const { return (word & 4) != 0; }
  void set_axis(unsigned a) { word = (word & ~3u) + a; }
  unsigned axis() const { return word & 3; }
  void set_leaf() { word = (word & ~3u) + 3; }
  bool leaf() const { return (word & 3) == 3; }
};

template <class P>
struct KD {
  std::vector<P> pts;
  std::vector<NodeK> nodes;

  explicit KD(std::vector<P> &&v) : pts(std::move(v)) {
    nodes.assign(pts.size(), NodeK());
    if (!pts.empty()) build(0, pts.size());
  }

  void build(size_t b, size_t e) {
    if (b + 1 == e) {
      nodes[b].set_leaf();
      return;
    }
    BBoxD bb;
    for (size_t i = b; i < e; ++i) bb.unite(V3{(double)pts[i][0], (double)pts[i][1], (double)pts[i][2]});
    int axis = 0;
    if (bb.mx[1] - bb.mn[1] > bb.mx[axis] - bb.mn[axis]) axis = 1;
    if (bb.mx[2] - bb.mn[2] > bb.mx[axis] - bb.mn[axis]) axis = 2;
    size_t split = (b + e) / 2;
    std::nth_element(pts.begin() + b, pts.begin() + split, pts.begin() + e,
                     [axis](const P &p1, const P &p2) { return p1[axis] < p2[axis]; });
    std::swap(pts[b], pts[split]);
    nodes[b].set_axis((unsigned)axis);
    nodes[b].split = (float)pts[b][axis];
    if (b < split) {
      nodes[b].set_has_left(true);
      build(b + 1, split + 1);
    } else {
      nodes[b].set_has_left(false);
    }
    if (split + 1 < e) {
      nodes[b].set_right((uint32_t)split + 1);
      build(split + 1, e);
    } else {
      nodes[b].set_right((1u << 29) - 1);
    }
  }

  // cnt (optional): [0] calls of _Lookup (nodes visited), [1] points inside the radius (those the
  // processor is called for, KDTree.h:549-551).
  template <class Proc>
  void lookup(size_t i, V3 q, Proc &proc, double &max_d2, uint32_t *cnt = nullptr) const {
    if (nodes.empty()) return;
    const NodeK &n = nodes[i];
    const P &p = pts[i];
    double dx = p[0] - q.x, dy = p[1] - q.y, dz = p[2] - q.z;
    double d2 = dx * dx + dy * dy + dz * dz;
    if (cnt) ++cnt[0];
    if (d2 < max_d2) {
      if (cnt) ++cnt[1];
      proc(i, d2, max_d2);
    }
    if (!n.leaf()) {
      unsigned a = n.axis();
      double qa = q[(int)a];
      double pd2 = (qa - n.split) * (qa - n.split);
      if (qa <= n.split) {
        if (n.has_left()) lookup(i + 1, q, proc, max_d2, cnt);
        if (pd2 < max_d2 && n.right() < nodes.size()) lookup(n.right(), q, proc, max_d2, cnt);
      } else {
        if (n.right() < nodes.size()) lookup(n.right(), q, proc, max_d2, cnt);
        if (pd2 < max_d2 && n.has_left()) lookup(i + 1, q, proc, max_d2, cnt);
      }
    }
  }
};

struct NP {
  uint32_t pos;
  double d2;
  bool operator<(const NP &o) const { return d2 == o.d2 ? (pos < o.pos) : (d2 < o.d2); }
};

template <class Filter>
struct KnnProc {
  NP *arr;
  size_t k, size = 0;
  Filter filt;
  void operator()(size_t i, double d2, double &max_d2) {
    if (!filt(i)) return;
    if (size < k) {
      arr[size++] = NP{(uint32_t)i, d2};
      if (size == k) {
        std::make_heap(arr, arr + size);
        max_d2 = arr[0].d2;
      }
    } else {
      std::pop_heap(arr, arr + size);
      arr[k - 1] = NP{(uint32_t)i, d2};
      std::push_heap(arr, arr + size);
      max_d2 = arr[0].d2;
    }
  }
};

template <class Filter>
struct NearestProc {
  int64_t best = -1;
  Filter filt;
  void operator()(size_t i, double d2, double &max_d2) {
    if (!filt(i)) return;
    best = (int64_t)i;
    max_d2 = d2;
  }
};

struct PtD {
  double c[3];
  uint32_t id;
  double operator[](int i) const { return c[i]; }
};
struct PtF {
  float c[3];
  uint32_t id;
  float operator[](int i) const { return c[i]; }
};

struct Photon {
  float p[3];
  float w[3];
  uint16_t dir, normal;
  uint32_t id;
  float operator[](int i) const { return p[i]; }
};

template <class P>
struct CustomFilter {
  const KD<P> *t;
  bool operator()(size_t i) const {
    const P &p = t->pts[i];
    int sum = (int)((double)p[0] + (double)p[1] + (double)p[2]);
    return (sum % 2) == 1;
  }
};
struct NoFilter {
  bool operator()(size_t) const { return true; }
};

// PhotonInternalTypes.h:113-142
struct PhotonFilter {
  const KD<Photon> *t;
  V3 sp, sn;
  double cos_t, sin_t;
  PhotonFilter(const KD<Photon> *tree, V3 p, V3 n, double c) : t(tree), sp(p), sn(n), cos_t(c) {
    sin_t = std::sqrt(1.0 - cos_t * cos_t);
  }
  bool operator()(size_t i) const {
    const Photon &ph = t->pts[i];
    V3 to = sub(V3{(double)ph.p[0], (double)ph.p[1], (double)ph.p[2]}, sp);
    return dot(cdir(ph.normal), sn) > cos_t && std::fabs(dot(to, sn)) < sin_t * std::sqrt(len_sqr(to));
  }
};

static const double MAX_NORMAL_DEVIATION_COS = 0.87;

struct PhotonMap {
  KD<Photon> tree;
  explicit PhotonMap(std::vector<Photon> &&v) : tree(std::move(v)) {}

  // GetNearestPoint with PhotonFilter (KDTree.h:484-503, NearestPointProc :293-323); the
  // IrradiancePhotonFilter of PhotonInternalTypes.h:153-168 makes the same test. Tree position or -1.
  int64_t nearest(V3 p, V3 n, double r) const {
    if (tree.pts.empty()) return -1;
    NearestProc<PhotonFilter> proc{-1, PhotonFilter(&tree, p, n, MAX_NORMAL_DEVIATION_COS)};
    double max_d2 = r * r;
    tree.lookup(0, p, proc, max_d2);
    return proc.best;
  }

  // Returns photons found; arr holds them in the processor's final order. cnt: see KD::lookup.
  size_t knn(V3 p, V3 n, uint32_t k, double r, NP *arr, uint32_t *cnt = nullptr) const {
    if (k == 0 || tree.pts.empty()) return 0;
    KnnProc<PhotonFilter> proc{arr, k, 0, PhotonFilter(&tree, p, n, MAX_NORMAL_DEVIATION_COS)};
    double max_d2 = r * r;
    tree.lookup(0, p, proc, max_d2, cnt);
    return proc.size;
  }
};

template <class P>
static void kd_dump(const KD<P> &t, uint32_t *perm, float *split, uint32_t *word) {
  for (size_t i = 0; i < t.pts.size(); ++i) {
    if (perm) perm[i] = t.pts[i].id;
    if (split) split[i] = t.nodes[i].split;
    if (word) word[i] = t.nodes[i].word;
  }
}

template <class P>
static void kd_knn(const KD<P> &t, uint64_t nq, const double *q, const double *radius, const uint32_t *k,
                   uint32_t kmax, int filter, uint32_t *ids, double *d2, uint32_t *count, int nthreads) {
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<NP> arr(kmax > 0 ? kmax : 1);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
      V3 qp{q[3 * i], q[3 * i + 1], q[3 * i + 2]};
      uint32_t kk = k[i];
      size_t found = 0;
      if (kk > 0 && !t.pts.empty()) {
        double max_d2 = radius[i] * radius[i];
        if (filter == 1) {
          KnnProc<CustomFilter<P>> proc{arr.data(), kk, 0, CustomFilter<P>{&t}};
          t.lookup(0, qp, proc, max_d2);
          found = proc.size;
        } else {
          KnnProc<NoFilter> proc{arr.data(), kk, 0, NoFilter{}};
          t.lookup(0, qp, proc, max_d2);
          found = proc.size;
        }
      }
      count[i] = (uint32_t)found;
      for (size_t j = 0; j < found; ++j) {
        ids[i * kmax + j] = t.pts[arr[j].pos].id;
        d2[i * kmax + j] = arr[j].d2;
      }
    }
  }
}

template <class P>
static void kd_nearest(const KD<P> &t, uint64_t nq, const double *q, const double *radius, int filter, uint32_t *id,
                       int nthreads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)nq; ++i) {
    V3 qp{q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    double max_d2 = radius[i] * radius[i];
    int64_t best = -1;
    if (filter == 1) {
      NearestProc<CustomFilter<P>> proc{-1, CustomFilter<P>{&t}};
      t.lookup(0, qp, proc, max_d2);
      best = proc.best;
    } else {
      NearestProc<NoFilter> proc{-1, NoFilter{}};
      t.lookup(0, qp, proc, max_d2);
      best = proc.best;
    }
    id[i] = best < 0 ? 0xffffffffu : t.pts[(size_t)best].id;
  }
}

}  // namespace orc

using namespace orc;

struct orc_kdtree {
  int kind;  // 0 double points, 1 float points
  KD<PtD> *d = nullptr;
  KD<PtF> *f = nullptr;
};
struct orc_photon_map {
  PhotonMap *m;
};

extern "C" {

orc_kdtree *orc_kdtree_create_d(uint64_t n, const double *pts) {
  std::vector<PtD> v(n);
  for (uint64_t i = 0; i < n; ++i) v[i] = PtD{{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, (uint32_t)i};
  orc_kdtree *h = new orc_kdtree();
  h->kind = 0;
  h->d = new KD<PtD>(std::move(v));
  return h;
}

orc_kdtree *orc_kdtree_create_f(uint64_t n, const float *pts) {
  std::vector<PtF> v(n);
  for (uint64_t i = 0; i < n; ++i) v[i] = PtF{{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, (uint32_t)i};
  orc_kdtree *h = new orc_kdtree();
  h->kind = 1;
  h->f = new KD<PtF>(std::move(v));
  return h;
}

void orc_kdtree_destroy(orc_kdtree *h) {
  delete h->d;
  delete h->f;
  delete h;
}

void orc_kdtree_dump(const orc_kdtree *h, uint32_t *perm, float *split, uint32_t *word) {
  if (h->kind == 0) kd_dump(*h->d, perm, split, word);
  else kd_dump(*h->f, perm, split, word);
}

void orc_kdtree_knn(const orc_kdtree *h, uint64_t nq, const double *q, const double *radius, const uint32_t *k,
                    uint32_t kmax, int filter, uint32_t *ids, double *d2, uint32_t *count, int nthreads) {
  if (h->kind == 0) kd_knn(*h->d, nq, q, radius, k, kmax, filter, ids, d2, count, nthreads);
  else kd_knn(*h->f, nq, q, radius, k, kmax, filter, ids, d2, count, nthreads);
}

void orc_kdtree_nearest(const orc_kdtree *h, uint64_t nq, const double *q, const double *radius, int filter,
                        uint32_t *id, int nthreads) {
  if (h->kind == 0) kd_nearest(*h->d, nq, q, radius, filter, id, nthreads);
  else kd_nearest(*h->f, nq, q, radius, filter, id, nthreads);
}

orc_photon_map *orc_photon_map_create(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                                      const uint16_t *normal) {
  cdir_init();
  std::vector<Photon> v(n);
  for (uint64_t i = 0; i < n; ++i) {
    Photon &p = v[i];
    for (int c = 0; c < 3; ++c) {
      p.p[c] = pos[3 * i + c];
      p.w[c] = weight[3 * i + c];
    }
    p.dir = dir[i];
    p.normal = normal[i];
    p.id = (uint32_t)i;
  }
  orc_photon_map *h = new orc_photon_map();
  h->m = new PhotonMap(std::move(v));
  return h;
}

void orc_photon_map_destroy(orc_photon_map *h) {
  delete h->m;
  delete h;
}

void orc_photon_map_dump(const orc_photon_map *h, uint32_t *perm, float *split, uint32_t *word) {
  kd_dump(h->m->tree, perm, split, word);
}

// PhotonLTEIntegrator.cpp:127-169
void orc_photon_irradiance(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                           double r, uint64_t paths, float *e_ext, float *e_int, uint32_t *ids, double *d2,
                           uint32_t *count, int nthreads) {
  const PhotonMap &m = *h->m;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<NP> arr(k > 0 ? k : 1);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
      V3 qp{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, qn{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
      size_t found = m.knn(qp, qn, k, r, arr.data());
      if (count) count[i] = (uint32_t)found;
      if (ids)
        for (size_t j = 0; j < found; ++j) {
          ids[i * k + j] = m.tree.pts[arr[j].pos].id;
          d2[i * k + j] = arr[j].d2;
        }
      for (int c = 0; c < 3; ++c) e_ext[3 * i + c] = e_int[3 * i + c] = 0.0f;
      if (found == 0) continue;
      double max_d2 = 0.0;
      double ext[3] = {0, 0, 0}, in[3] = {0, 0, 0};
      for (size_t j = 0; j < found; ++j) {
        const Photon &ph = m.tree.pts[arr[j].pos];
        V3 pp{(double)ph.p[0], (double)ph.p[1], (double)ph.p[2]};
        double t = len_sqr(sub(pp, qp));
        if (t > max_d2) max_d2 = t;
        double *acc = dot(cdir(ph.dir), qn) > 0.0 ? ext : in;
        for (int c = 0; c < 3; ++c) acc[c] += (double)ph.w[c];
      }
      if (found < k || max_d2 == 0.0) max_d2 = r * r;
      else max_d2 *= (double)found / ((double)found - 0.5);
      double inv = 1.0 / (PI_ * (double)paths * max_d2);
      for (int c = 0; c < 3; ++c) {
        e_ext[3 * i + c] = (float)(ext[c] * inv);
        e_int[3 * i + c] = (float)(in[c] * inv);
      }
    }
  }
}

// The work of each lookup of orc_photon_irradiance: nodes visited (calls of KDTree::_Lookup,
// KDTree.h:535-576), photons inside the radius (KDTree.h:549) and the final set size.
void orc_photon_lookup_counts(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                              double r, uint32_t *counts, int nthreads) {
  const PhotonMap &m = *h->m;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<NP> arr(k > 0 ? k : 1);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
      V3 qp{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, qn{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
      uint32_t cnt[2] = {0, 0};
      size_t found = m.knn(qp, qn, k, r, arr.data(), cnt);
      counts[3 * i] = cnt[0];
      counts[3 * i + 1] = cnt[1];
      counts[3 * i + 2] = (uint32_t)found;
    }
  }
}

// PhotonLTEIntegrator.cpp:579-636 with BSDF::Evaluate of a Lambertian (BSDF.cpp:22-45, Lambertian.cpp:25-30)
void orc_photon_caustic(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, const double *wo,
                        uint32_t k, double r, uint64_t paths, double rho, double *radiance, int nthreads) {
  const PhotonMap &m = *h->m;
  double f = rho * INV_PI_;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<NP> arr(k > 0 ? k : 1);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
      V3 qp{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, qn{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
      V3 w{wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]};
      size_t found = m.knn(qp, qn, k, r, arr.data());
      for (int c = 0; c < 3; ++c) radiance[3 * i + c] = 0.0;
      if (found == 0) continue;
      double max_d2 = 0.0;
      double rad[3] = {0, 0, 0};
      for (size_t j = 0; j < found; ++j) {
        const Photon &ph = m.tree.pts[arr[j].pos];
        V3 pp{(double)ph.p[0], (double)ph.p[1], (double)ph.p[2]};
        V3 pdir = cdir(ph.dir);
        double t = len_sqr(sub(pp, qp));
        if (t > max_d2) max_d2 = t;
        double tmp = 1.0 - len_sqr(sub(pp, qp)) / max_d2;
        double kernel = 3.0 * INV_PI_ * tmp * tmp;
        double fv = (dot(pdir, qn) * dot(w, qn) > 0) ? f : 0.0;
        for (int c = 0; c < 3; ++c) rad[c] += fv * (double)ph.w[c] * kernel;
      }
      if (found < k || max_d2 == 0.0) max_d2 = r * r;
      else max_d2 *= (double)found / ((double)found - 0.5);
      double div = (double)paths * max_d2;
      double inv = 1.0 / div;
      for (int c = 0; c < 3; ++c) radiance[3 * i + c] = rad[c] * inv;
    }
  }
}

// Cone filter over the same k-nearest set (not in the reference: parity of this variant rests
// on this file alone).
void orc_photon_cone(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, uint32_t k,
                     double r, uint64_t paths, double kc, float *e_ext, int nthreads) {
  const PhotonMap &m = *h->m;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    std::vector<NP> arr(k > 0 ? k : 1);
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
      V3 qp{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, qn{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
      size_t found = m.knn(qp, qn, k, r, arr.data());
      for (int c = 0; c < 3; ++c) e_ext[3 * i + c] = 0.0f;
      if (found == 0) continue;
      double max_d2 = 0.0;
      for (size_t j = 0; j < found; ++j) max_d2 = std::max(max_d2, arr[j].d2);
      double rmax2 = (found < k || max_d2 == 0.0) ? r * r : max_d2;
      double rmax = std::sqrt(rmax2);
      double ext[3] = {0, 0, 0};
      for (size_t j = 0; j < found; ++j) {
        const Photon &ph = m.tree.pts[arr[j].pos];
        if (!(dot(cdir(ph.dir), qn) > 0.0)) continue;
        double wgt = 1.0 - std::sqrt(arr[j].d2) / (kc * rmax);
        for (int c = 0; c < 3; ++c) ext[c] += wgt * (double)ph.w[c];
      }
      double norm = (1.0 - 2.0 / (3.0 * kc)) * PI_ * (double)paths * rmax2;
      for (int c = 0; c < 3; ++c) e_ext[3 * i + c] = (float)(ext[c] / norm);
    }
  }
}

void orc_photon_nearest(const orc_photon_map *h, uint64_t nq, const double *p, const double *n, double r, uint32_t *id,
                        int nthreads) {
  const PhotonMap &m = *h->m;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t i = 0; i < (int64_t)nq; ++i) {
    V3 qp{p[3 * i], p[3 * i + 1], p[3 * i + 2]}, qn{n[3 * i], n[3 * i + 1], n[3 * i + 2]};
    const int64_t b = m.nearest(qp, qn, r);
    id[i] = b < 0 ? 0xffffffffu : m.tree.pts[(size_t)b].id;
  }
}

void orc_photon_map_points(const orc_photon_map *h, float *pos, uint16_t *dir, uint16_t *normal) {
  const PhotonMap &m = *h->m;
  for (size_t i = 0; i < m.tree.pts.size(); ++i) {
    const Photon &ph = m.tree.pts[i];
    for (int c = 0; c < 3; ++c) pos[3 * i + c] = ph.p[c];
    if (dir) dir[i] = ph.dir;
    if (normal) normal[i] = ph.normal;
  }
}

uint64_t orc_photon_map_size(const orc_photon_map *h) { return h->m->tree.pts.size(); }

}  // extern "C"

// ------------------------------------------------------------------ irradiance photon map

struct orc_irr_map {
  orc_photon_map *map = nullptr;  // the irradiance photons (weight = external irradiance)
  std::vector<float> pos, e_ext, e_int;
  std::vector<uint16_t> normal;
  double max_dist = DBL_INF_;
  ~orc_irr_map() {
    if (map) orc_photon_map_destroy(map);
  }
};

static void irr_build_tree(orc_irr_map *m) {
  const uint64_t n = m->normal.size();
  std::vector<uint16_t> dir(n, 0);
  m->map = orc_photon_map_create(n, m->pos.data(), m->e_ext.data(), dir.data(), m->normal.data());
}

extern "C" {

// PhotonLTEIntegrator::_ConstructIrradiancePhotonMap (PhotonLTEIntegrator.cpp:171-245) with the
// per-photon estimate of IrradiancePhotonProcess (PhotonInternalTypes.h:269-288).
orc_irr_map *orc_irr_map_create(const orc_photon_map *caustic, const orc_photon_map *direct,
                                const orc_photon_map *indirect, uint64_t caustic_paths, uint64_t direct_paths,
                                uint64_t indirect_paths, double scene_area, int nthreads) {
  if (!indirect || !direct) return nullptr;
  cdir_init();
  const std::vector<Photon> &ind = indirect->m->tree.pts;  // GetAllPoints: tree order
  orc_irr_map *m = new orc_irr_map();
  for (size_t i = 0; i < ind.size(); i += 10) {
    for (int c = 0; c < 3; ++c) m->pos.push_back(ind[i].p[c]);
    m->normal.push_back(ind[i].normal);
  }
  const uint64_t n = m->normal.size();
  m->e_ext.assign(3 * n, 0.0f);
  m->e_int.assign(3 * n, 0.0f);
  if (n == 0) return m;  // the reference leaves the map null: no lookups find anything
  const size_t nd = direct->m->tree.pts.size(), ni = ind.size();
  const double direct_area = nd == 0 ? 0.0 : scene_area / (double)nd;
  const double r_direct = std::sqrt(direct_area * 100.0 * INV_PI_);
  const double indirect_area = ni == 0 ? 0.0 : scene_area / (double)ni;
  const double r_indirect = std::sqrt(indirect_area * 100.0 * INV_PI_);
  const double r_caustic = std::max(r_direct, r_indirect);
  std::vector<double> pts(3 * n), nrm(3 * n);
  for (uint64_t i = 0; i < n; ++i) {
    const double *dv = g_cdir[m->normal[i]];
    for (int c = 0; c < 3; ++c) {
      pts[3 * i + c] = (double)m->pos[3 * i + c];
      nrm[3 * i + c] = dv[c];
    }
  }
  std::vector<float> ce(3 * n, 0.0f), ci(3 * n, 0.0f), de(3 * n), di(3 * n), ie(3 * n), ii(3 * n);
  if (caustic)
    orc_photon_irradiance(caustic, n, pts.data(), nrm.data(), 100, r_caustic, caustic_paths, ce.data(), ci.data(),
                          nullptr, nullptr, nullptr, nthreads);
  orc_photon_irradiance(direct, n, pts.data(), nrm.data(), 100, r_direct, direct_paths, de.data(), di.data(), nullptr,
                        nullptr, nullptr, nthreads);
  orc_photon_irradiance(indirect, n, pts.data(), nrm.data(), 100, r_indirect, indirect_paths, ie.data(), ii.data(),
                        nullptr, nullptr, nullptr, nthreads);
  for (uint64_t j = 0; j < 3 * n; ++j) {
    m->e_ext[j] = (ce[j] + de[j]) + ie[j];
    m->e_int[j] = (ci[j] + di[j]) + ii[j];
  }
  irr_build_tree(m);
  // The 0.1 % percentile of the nearest-irradiance-photon distances of every 11th indirect photon.
  const size_t step = 11, nq = (ni + step - 1) / step;
  std::vector<double> qp(3 * nq), qn(3 * nq), sq(nq, 0.0);
  for (size_t i = 0; i < nq; ++i) {
    const Photon &ph = ind[i * step];
    const double *dv = g_cdir[ph.normal];
    for (int c = 0; c < 3; ++c) {
      qp[3 * i + c] = (double)ph.p[c];
      qn[3 * i + c] = dv[c];
    }
  }
  std::vector<uint32_t> id(nq);
  orc_photon_nearest(m->map, nq, qp.data(), qn.data(), DBL_INF_, id.data(), nthreads);
  for (size_t i = 0; i < nq; ++i) {
    if (id[i] == 0xffffffffu) continue;
    V3 a{qp[3 * i], qp[3 * i + 1], qp[3 * i + 2]};
    V3 b{(double)m->pos[3 * id[i]], (double)m->pos[3 * id[i] + 1], (double)m->pos[3 * id[i] + 2]};
    sq[i] = len_sqr(sub(a, b));
  }
  const size_t index = sq.size() <= 1 ? 0 : std::max(sq.size() / 1000, (size_t)1);
  if (!sq.empty()) {
    std::nth_element(sq.begin(), sq.begin() + index, sq.end(), std::greater<double>());
    m->max_dist = sq[index] == 0.0 ? DBL_INF_ : std::sqrt(sq[index]);
  }
  return m;
}

// An irradiance photon map from explicit photons (creation order = caller order).
orc_irr_map *orc_irr_map_create_from(uint64_t n, const float *pos, const uint16_t *normal, const float *e_ext,
                                     const float *e_int, double max_dist) {
  cdir_init();
  orc_irr_map *m = new orc_irr_map();
  m->pos.assign(pos, pos + 3 * n);
  m->normal.assign(normal, normal + n);
  m->e_ext.assign(e_ext, e_ext + 3 * n);
  m->e_int.assign(e_int, e_int + 3 * n);
  m->max_dist = max_dist;
  irr_build_tree(m);
  return m;
}

void orc_irr_map_destroy(orc_irr_map *m) { delete m; }

void orc_irr_map_info(const orc_irr_map *m, uint64_t *n, double *max_dist) {
  if (n) *n = m->normal.size();
  if (max_dist) *max_dist = m->max_dist;
}

void orc_irr_map_export(const orc_irr_map *m, float *pos, uint16_t *normal, float *e_ext, float *e_int) {
  const size_t n = m->normal.size();
  if (pos) std::copy(m->pos.begin(), m->pos.begin() + 3 * n, pos);
  if (normal) std::copy(m->normal.begin(), m->normal.end(), normal);
  if (e_ext) std::copy(m->e_ext.begin(), m->e_ext.end(), e_ext);
  if (e_int) std::copy(m->e_int.begin(), m->e_int.end(), e_int);
}

// The nearest irradiance photon of each query with IrradiancePhotonFilter within r (creation index
// or 0xFFFFFFFF).
void orc_irr_map_nearest(const orc_irr_map *m, uint64_t nq, const double *p, const double *n, double r, uint32_t *id,
                         int nthreads) {
  if (!m->map) {
    for (uint64_t i = 0; i < nq; ++i) id[i] = 0xffffffffu;
    return;
  }
  orc_photon_nearest(m->map, nq, p, n, r, id, nthreads);
}

uint16_t orc_cdir_encode(double x, double y, double z) { return cdir_encode(V3{x, y, z}); }

void orc_cdir_table(double *out) {
  cdir_init();
  for (int i = 0; i < (1 << 16); ++i)
    for (int c = 0; c < 3; ++c) out[3 * i + c] = g_cdir[i][c];
}

void orc_mt19937(uint32_t seed, uint64_t n, uint32_t *out) {
  std::mt19937 g(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = (uint32_t)g();
}

}  // extern "C"
