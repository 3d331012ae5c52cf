// ORACLE — TEST INFRASTRUCTURE ONLY.
// Part of the CPU restatement of Skwarka's ray/photon hot path that the tests, smoke() and the
// CPU-baseline leg of bench.py use as the checker. Nothing in the product library links or calls it.
//
// Value semantics of the reference's L0 math types, restated:
//   Vector3D dot product  = (x*x' + y*y') + z*z'          Source/RayLibs/Math/Vector3D.h (operator*)
//   Vector3D cross product = (y z' - z y', z x' - x z', x y' - y x')   Vector3D.h (operator^)
//   Normalize = multiply by 1.0/sqrt(len^2) (operator/= takes the reciprocal)   Vector3D.h:209-270
//   BBox3D default = (+inf, -inf); Unite uses std::min/std::max; Area = 2(xy+xz+yz) in T
//                                                       Source/RayLibs/Math/BBox3D.h:121-141,188-200
// Build with -ffp-contract=off: the reference is an SSE2 /fp:precise build without FMA.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>

namespace orc {

constexpr double DBL_EPS_ = std::numeric_limits<double>::epsilon();
constexpr double DBL_INF_ = std::numeric_limits<double>::infinity();
constexpr double PI_ = 3.14159265358979323846;
constexpr double INV_PI_ = 0.31830988618379067154;

// std::min / std::max exactly as libstdc++ and MSVC define them (the first argument wins ties
// and NaN comparisons).
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }

struct V3 {
  double x, y, z;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double len_sqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline bool normalize(V3 &a) {
  double l2 = len_sqr(a);
  if (l2 > 0.0) {
    double inv = 1.0 / std::sqrt(l2);
    a = mul(a, inv);
    return true;
  }
  return false;
}

struct F3 {
  float x, y, z;
  float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 to_d(F3 p) { return {(double)p.x, (double)p.y, (double)p.z}; }

struct BBoxF {
  float mn[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                 std::numeric_limits<float>::infinity()};
  float mx[3] = {-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                 -std::numeric_limits<float>::infinity()};
  void unite(F3 p) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = smin(mn[i], p[i]);
      mx[i] = smax(mx[i], p[i]);
    }
  }
  void unite(const BBoxF &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = smin(mn[i], b.mn[i]);
      mx[i] = smax(mx[i], b.mx[i]);
    }
  }
  float area() const {
    float x = std::fabs(mx[0] - mn[0]);
    float y = std::fabs(mx[1] - mn[1]);
    float z = std::fabs(mx[2] - mn[2]);
    return 2.0f * (x * y + x * z + y * z);
  }
};

struct BBoxD {
  double mn[3] = {DBL_INF_, DBL_INF_, DBL_INF_};
  double mx[3] = {-DBL_INF_, -DBL_INF_, -DBL_INF_};
  void unite(const BBoxF &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = smin(mn[i], (double)b.mn[i]);
      mx[i] = smax(mx[i], (double)b.mx[i]);
    }
  }
  void unite(const BBoxD &b) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = smin(mn[i], b.mn[i]);
      mx[i] = smax(mx[i], b.mx[i]);
    }
  }
  void unite(V3 p) {
    for (int i = 0; i < 3; ++i) {
      mn[i] = smin(mn[i], p[i]);
      mx[i] = smax(mx[i], p[i]);
    }
  }
  double area() const {
    double x = std::fabs(mx[0] - mn[0]);
    double y = std::fabs(mx[1] - mn[1]);
    double z = std::fabs(mx[2] - mn[2]);
    return 2.0 * (x * y + x * z + y * z);
  }
};

// Transform as a (matrix, inverse) pair of row-major 4x4 doubles.
// Source/RayLibs/Math/Transform.h:197-300 (operator() on points, vectors and rays; Inverted swaps).
struct Xf {
  double m[4][4];
  double inv[4][4];
  bool identity;
};
inline V3 xf_point(const double m[4][4], V3 p) {
  double xp = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
  double yp = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
  double zp = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
  double w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
  V3 r{xp, yp, zp};
  if (w != 1.0) r = mul(r, 1.0 / w);
  return r;
}
// Point3D_f transform: double arithmetic, each coordinate cast to float; the w division happens
// on the float point (Point3D<float>::operator/ multiplies by the double reciprocal and casts).
inline F3 xf_point_f(const double m[4][4], F3 p) {
  double x = p.x, y = p.y, z = p.z;
  float xp = (float)(m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3]);
  float yp = (float)(m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3]);
  float zp = (float)(m[2][0] * x + m[2][1] * y + m[2][2] * z + m[2][3]);
  float wp = (float)(m[3][0] * x + m[3][1] * y + m[3][2] * z + m[3][3]);
  if (wp == 1.0f) return {xp, yp, zp};
  double inv = 1.0 / (double)wp;
  return {(float)(xp * inv), (float)(yp * inv), (float)(zp * inv)};
}
inline V3 xf_vector(const double m[4][4], V3 v) {
  return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
          m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}
// TransformNormal uses the transposed inverse and normalises (Transform.h:281-290).
inline V3 xf_normal(const double inv[4][4], V3 n) {
  V3 r{inv[0][0] * n.x + inv[1][0] * n.y + inv[2][0] * n.z, inv[0][1] * n.x + inv[1][1] * n.y + inv[2][1] * n.z,
       inv[0][2] * n.x + inv[1][2] * n.y + inv[2][2] * n.z};
  normalize(r);
  return r;
}

struct Ray {
  V3 o, d;
  double tmin, tmax;
};

}  // namespace orc
