// PBRT (v2) scene import for the path: the subset of the reference's importer
// (CADImport/SceneImporters/PbrtSceneImporter: PbrtSceneImporter_main.cpp, _hooks.cpp, Factories/)
// that feeds the triangle scene, its matte / plastic materials, diffuse area lights and the
// perspective camera. Supported commands: Identity, Translate, Scale, Rotate, LookAt, Transform,
// ConcatTransform, CoordinateSystem, CoordSysTransform, AttributeBegin / End, TransformBegin /
// End, WorldBegin / End, ReverseOrientation, Camera "perspective", Film, Shape "trianglemesh",
// Material "matte" / "plastic", MakeNamedMaterial, NamedMaterial, AreaLightSource "diffuse",
// ObjectBegin / End, ObjectInstance, Include. Every other command, shape, material or parameter
// is skipped with a warning (skw_pbrt_warning).
//
// Semantics kept from the reference: numbers of transform commands and float parameters are
// parsed as float (std::stof), points, vectors and normals as double (std::stod); the current
// transform is post-multiplied (Transform * Make...); MakeLookAt is the inverse of the camera
// frame (Math/Transform.cpp:138-165); matrices are inverted by Gauss-Jordan elimination with
// full pivoting (Math/Matrix4x4.h:240-300); static shapes are placed in world space (double math,
// then float vertices) with an identity primitive transform; instances share the object's meshes
// under the current transform at ObjectInstance; the camera-to-world transform is the inverse of
// the current transform at Camera, mirrored in x when that transform inverts orientation
// (_pbrtCamera); the x view angle is 2 atan(tan(fov / 2) max(1, frame aspect)).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace skw;

namespace {

struct M4 {
  double m[4][4];
  static M4 identity() {
    M4 r;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) r.m[i][j] = i == j ? 1.0 : 0.0;
    return r;
  }
};

M4 mul(const M4 &a, const M4 &b) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] + a.m[i][3] * b.m[3][j];
  return r;
}

// Gauss-Jordan elimination with full pivoting; false when singular.
bool invert(const M4 &in, M4 &out) {
  int indxc[4], indxr[4], ipiv[4] = {0, 0, 0, 0};
  double a[4][4];
  std::memcpy(a, in.m, sizeof(a));
  for (int i = 0; i < 4; ++i) {
    int irow = -1, icol = -1;
    double big = 0.0;
    for (int j = 0; j < 4; ++j) {
      if (ipiv[j] == 1) continue;
      for (int k = 0; k < 4; ++k) {
        if (ipiv[k] == 0) {
          if (std::fabs(a[j][k]) >= big) {
            big = std::fabs(a[j][k]);
            irow = j;
            icol = k;
          }
        } else if (ipiv[k] > 1) {
          return false;
        }
      }
    }
    if (irow < 0) return false;
    ++ipiv[icol];
    if (irow != icol)
      for (int k = 0; k < 4; ++k) std::swap(a[irow][k], a[icol][k]);
    indxr[i] = irow;
    indxc[i] = icol;
    if (a[icol][icol] == 0.0) return false;
    const double pivinv = 1.0 / a[icol][icol];
    a[icol][icol] = 1.0;
    for (int j = 0; j < 4; ++j) a[icol][j] *= pivinv;
    for (int j = 0; j < 4; ++j) {
      if (j == icol) continue;
      const double save = a[j][icol];
      a[j][icol] = 0.0;
      for (int k = 0; k < 4; ++k) a[j][k] -= a[icol][k] * save;
    }
  }
  for (int j = 3; j >= 0; --j)
    if (indxr[j] != indxc[j])
      for (int k = 0; k < 4; ++k) std::swap(a[k][indxr[j]], a[k][indxc[j]]);
  std::memcpy(out.m, a, sizeof(a));
  return true;
}

// A transform and its inverse (Math/Transform.h).
struct Xf {
  M4 m = M4::identity(), inv = M4::identity();
  static Xf from(const M4 &m) {
    Xf x;
    x.m = m;
    if (!invert(m, x.inv)) x.inv = M4::identity();
    return x;
  }
  Xf inverted() const {
    Xf x;
    x.m = inv;
    x.inv = m;
    return x;
  }
  Xf operator*(const Xf &o) const {
    Xf x;
    x.m = mul(m, o.m);
    x.inv = mul(o.inv, inv);
    return x;
  }
  bool inverts_orientation() const {
    const auto &v = m.m;
    const double det = ((v[0][0] * (v[1][1] * v[2][2] - v[1][2] * v[2][1])) -
                        (v[0][1] * (v[1][0] * v[2][2] - v[1][2] * v[2][0])) +
                        (v[0][2] * (v[1][0] * v[2][1] - v[1][1] * v[2][0])));
    return det < 0.0;
  }
  void point(const double p[3], double o[3]) const {
    const auto &v = m.m;
    double x = v[0][0] * p[0] + v[0][1] * p[1] + v[0][2] * p[2] + v[0][3];
    double y = v[1][0] * p[0] + v[1][1] * p[1] + v[1][2] * p[2] + v[1][3];
    double z = v[2][0] * p[0] + v[2][1] * p[1] + v[2][2] * p[2] + v[2][3];
    const double w = v[3][0] * p[0] + v[3][1] * p[1] + v[3][2] * p[2] + v[3][3];
    if (w != 1.0) {  // Point3D / w multiplies by the reciprocal
      const double iw = 1.0 / w;
      x *= iw;
      y *= iw;
      z *= iw;
    }
    o[0] = x;
    o[1] = y;
    o[2] = z;
  }
  void normal(const double n[3], double o[3]) const {
    const auto &v = inv.m;
    double x = v[0][0] * n[0] + v[1][0] * n[1] + v[2][0] * n[2];
    double y = v[0][1] * n[0] + v[1][1] * n[1] + v[2][1] * n[2];
    double z = v[0][2] * n[0] + v[1][2] * n[1] + v[2][2] * n[2];
    const double il = 1.0 / std::sqrt(x * x + y * y + z * z);
    o[0] = x * il;
    o[1] = y * il;
    o[2] = z * il;
  }
};

M4 make(double a, double b, double c, double d, double e, double f, double g, double h, double i, double j, double k,
        double l, double mm, double n, double o, double p) {
  M4 r;
  const double v[16] = {a, b, c, d, e, f, g, h, i, j, k, l, mm, n, o, p};
  for (int q = 0; q < 16; ++q) r.m[q / 4][q % 4] = v[q];
  return r;
}

void normalize(double v[3]) {
  const double il = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= il;
  v[1] *= il;
  v[2] *= il;
}

void cross(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

Xf make_rotation(double angle, double axis[3]) {
  normalize(axis);
  const double s = std::sin(angle), c = std::cos(angle);
  const double x = axis[0], y = axis[1], z = axis[2];
  return Xf::from(make(x * x + (1.0 - x * x) * c, x * y * (1.0 - c) - z * s, x * z * (1.0 - c) + y * s, 0.0,
                       x * y * (1.0 - c) + z * s, y * y + (1.0 - y * y) * c, y * z * (1.0 - c) - x * s, 0.0,
                       x * z * (1.0 - c) - y * s, y * z * (1.0 - c) + x * s, z * z + (1.0 - z * z) * c, 0.0, 0.0, 0.0,
                       0.0, 1.0));
}

Xf make_look_at(const double eye[3], double dir[3], const double up[3]) {
  normalize(dir);
  double right[3], nu[3];
  cross(dir, up, right);
  normalize(right);
  cross(right, dir, nu);
  const M4 m = make(right[0], -nu[0], dir[0], eye[0], right[1], -nu[1], dir[1], eye[1], right[2], -nu[2], dir[2],
                    eye[2], 0.0, 0.0, 0.0, 1.0);
  return Xf::from(m).inverted();
}

// ---------------------------------------------------------------- tokens and parameters

struct Tok {
  std::string s;
  bool quoted = false;
};

bool tokenize(const std::string &text, std::vector<Tok> &out, std::string &err) {
  size_t i = 0, n = text.size();
  while (i < n) {
    const char c = text[i];
    if (c == '#') {
      while (i < n && text[i] != '\n') ++i;
    } else if (std::isspace((unsigned char)c)) {
      ++i;
    } else if (c == '[' || c == ']') {
      out.push_back(Tok{std::string(1, c), false});
      ++i;
    } else if (c == '"') {
      const size_t j = text.find('"', i + 1);
      if (j == std::string::npos) {
        err = "unterminated string";
        return false;
      }
      out.push_back(Tok{text.substr(i + 1, j - i - 1), true});
      i = j + 1;
    } else {
      size_t j = i;
      while (j < n && !std::isspace((unsigned char)text[j]) && text[j] != '[' && text[j] != ']' && text[j] != '"' &&
             text[j] != '#')
        ++j;
      out.push_back(Tok{text.substr(i, j - i), false});
      i = j;
    }
  }
  return true;
}

struct Param {
  std::string type, name;
  std::vector<std::string> vals;
  bool used = false;
};

struct Params {
  std::vector<Param> list;
  Param *find(const char *name, const char *t1, const char *t2 = nullptr, const char *t3 = nullptr) {
    for (Param &p : list)
      if (p.name == name && (p.type == t1 || (t2 && p.type == t2) || (t3 && p.type == t3))) {
        p.used = true;
        return &p;
      }
    return nullptr;
  }
  float one_float(const char *name, float dflt) {
    Param *p = find(name, "float");
    return p && !p->vals.empty() ? std::stof(p->vals[0]) : dflt;
  }
  std::string one_string(const char *name, const std::string &dflt) {
    Param *p = find(name, "string");
    return p && !p->vals.empty() ? p->vals[0] : dflt;
  }
  int one_int(const char *name, int dflt) {
    Param *p = find(name, "integer");
    return p && !p->vals.empty() ? std::stoi(p->vals[0]) : dflt;
  }
  // rgb / color parameters (float values); `spectrum` and textures are not supported
  bool rgb(const char *name, double out[3]) {
    Param *p = find(name, "rgb", "color");
    if (!p || p->vals.size() < 3) return false;
    for (int c = 0; c < 3; ++c) out[c] = (double)std::stof(p->vals[c]);
    return true;
  }
};

struct Material {
  uint8_t type = 0;  // 0 matte, 1 plastic
  double kd[3] = {0.5, 0.5, 0.5}, ks[3] = {0.0, 0.0, 0.0}, rough = 0.1;
};

struct State {
  Xf ctm;
  Material mat;
  bool reverse = false;
  bool area = false;
  double L[3] = {1.0, 1.0, 1.0};
};

struct Mesh {
  std::vector<float> verts, normals;
  std::vector<uint32_t> tris;
  uint8_t flags = 0;
};

struct ObjShape {
  uint32_t mesh;
  Material mat;
};

}  // namespace

struct skw_pbrt_scene {
  std::vector<Mesh> meshes;
  std::vector<uint32_t> prim_mesh;
  std::vector<double> prim_xform;  // 32 per prim
  std::vector<uint8_t> mat_type;
  std::vector<double> kd, ks, rough;
  std::vector<uint32_t> light_prim;
  std::vector<double> light_radiance;
  std::vector<std::string> warnings;
  skw_camera cam;
  uint32_t width = 640, height = 480;
  bool has_camera = false;
  // flattened arrays handed out by skw_pbrt_descs
  std::vector<uint64_t> vo, to;
  std::vector<float> verts, normals;
  std::vector<uint32_t> tris;
  std::vector<uint8_t> flags;
};

namespace {

class Importer {
 public:
  explicit Importer(skw_pbrt_scene &out) : out_(out) {}

  bool run(const std::string &path, std::string &err, int depth = 0) {
    if (depth > 16) {
      err = "Include nesting deeper than 16";
      return false;
    }
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
      err = "cannot open " + path;
      return false;
    }
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    std::vector<Tok> toks;
    if (!tokenize(text, toks, err)) return false;
    const std::string dir = path.find('/') == std::string::npos ? std::string() : path.substr(0, path.rfind('/') + 1);
    size_t i = 0;
    try {
      while (i < toks.size()) {
        if (toks[i].quoted || toks[i].s == "[" || toks[i].s == "]") {
          err = "expected a command, found '" + toks[i].s + "'";
          return false;
        }
        const std::string cmd = toks[i++].s;
        if (!command(cmd, toks, i, dir, err, depth)) return false;
      }
    } catch (const std::exception &e) {
      err = std::string("bad number: ") + e.what();
      return false;
    }
    return true;
  }

 private:
  skw_pbrt_scene &out_;
  State st_;
  std::vector<State> stack_;
  std::vector<Xf> xf_stack_;
  std::map<std::string, Xf> coord_;
  std::map<std::string, Material> named_;
  std::map<std::string, std::vector<ObjShape>> objects_;
  std::vector<ObjShape> *cur_obj_ = nullptr;

  void warn(const std::string &w) { out_.warnings.push_back(w); }

  // numbers after a command: either n bare numbers or one bracketed list of n
  bool numbers(const std::vector<Tok> &t, size_t &i, int n, std::vector<float> &v, std::string &err) {
    v.clear();
    bool br = i < t.size() && t[i].s == "[" && !t[i].quoted;
    if (br) ++i;
    while ((int)v.size() < n && i < t.size() && !t[i].quoted && t[i].s != "]") v.push_back(std::stof(t[i++].s));
    if (br) {
      if (i >= t.size() || t[i].s != "]") {
        err = "expected ']'";
        return false;
      }
      ++i;
    }
    if ((int)v.size() != n) {
      err = "expected " + std::to_string(n) + " numbers";
      return false;
    }
    return true;
  }

  bool string_arg(const std::vector<Tok> &t, size_t &i, std::string &s, std::string &err) {
    if (i >= t.size() || !t[i].quoted) {
      err = "expected a quoted name";
      return false;
    }
    s = t[i++].s;
    return true;
  }

  bool params(const std::vector<Tok> &t, size_t &i, Params &ps, std::string &err) {
    while (i < t.size() && t[i].quoted) {
      const std::string decl = t[i++].s;
      const size_t sp = decl.find_first_of(" \t");
      if (sp == std::string::npos) {
        err = "bad parameter declaration \"" + decl + "\"";
        return false;
      }
      Param p;
      p.type = decl.substr(0, sp);
      p.name = decl.substr(decl.find_first_not_of(" \t", sp));
      if (i < t.size() && t[i].s == "[" && !t[i].quoted) {
        ++i;
        while (i < t.size() && !(t[i].s == "]" && !t[i].quoted)) p.vals.push_back(t[i++].s);
        if (i >= t.size()) {
          err = "expected ']'";
          return false;
        }
        ++i;
      } else if (i < t.size()) {
        p.vals.push_back(t[i++].s);
      }
      ps.list.push_back(p);
    }
    return true;
  }

  void unused(const char *what, Params &ps) {
    for (const Param &p : ps.list)
      if (!p.used) warn(std::string(what) + ": parameter \"" + p.type + " " + p.name + "\" ignored");
  }

  Material material(const std::string &name, Params &ps) {
    Material m;
    if (name == "matte") {
      m.type = 0;
      ps.rgb("Kd", m.kd);
      if (ps.one_float("sigma", 0.0f) != 0.0f) warn("matte: sigma > 0 (Oren-Nayar) is not supported; Lambertian used");
    } else if (name == "plastic") {
      m.type = 1;
      for (int c = 0; c < 3; ++c) m.kd[c] = m.ks[c] = 0.25;
      ps.rgb("Kd", m.kd);
      ps.rgb("Ks", m.ks);
      m.rough = (double)ps.one_float("roughness", 0.1f);
    } else {
      warn("material \"" + name + "\" is not supported; matte with Kd 0.5 used");
    }
    return m;
  }

  bool shape(const std::string &name, Params &ps) {
    if (name != "trianglemesh") {
      warn("shape \"" + name + "\" is not supported; skipped");
      return true;
    }
    Param *pi = ps.find("indices", "integer"), *pp = ps.find("P", "point");
    if (!pi || !pp) {
      warn("trianglemesh without indices or P; skipped");
      return true;
    }
    Param *pn = ps.find("N", "normal");
    const size_t np = pp->vals.size() / 3;
    Mesh m;
    std::vector<double> P(3 * np);
    for (size_t k = 0; k < 3 * np; ++k) P[k] = std::stod(pp->vals[k]);
    for (size_t k = 0; k < pi->vals.size(); ++k) {
      const long v = std::stol(pi->vals[k]);
      if (v < 0 || (size_t)v >= np) {
        warn("trianglemesh has an out-of-bounds vertex index; skipped");
        return true;
      }
    }
    for (size_t k = 0; k < np; ++k) {
      double o[3];
      st_.ctm.point(&P[3 * k], o);
      for (int c = 0; c < 3; ++c) m.verts.push_back((float)o[c]);
    }
    if (pn && pn->vals.size() / 3 == np) {
      for (size_t k = 0; k < np; ++k) {
        double nv[3] = {std::stod(pn->vals[3 * k]), std::stod(pn->vals[3 * k + 1]), std::stod(pn->vals[3 * k + 2])};
        normalize(nv);
        double o[3];
        st_.ctm.normal(nv, o);
        for (int c = 0; c < 3; ++c) m.normals.push_back((float)o[c]);
      }
      m.flags |= SKW_MESH_SHADING_NORMALS;
    } else if (pn) {
      warn("trianglemesh: number of N differs from P; normals ignored");
    }
    for (size_t k = 0; k + 2 < pi->vals.size(); k += 3)
      for (int c = 0; c < 3; ++c) m.tris.push_back((uint32_t)std::stol(pi->vals[k + c]));
    if (st_.reverse) m.flags |= SKW_MESH_INVERT_NORMALS;
    ps.find("uv", "float");
    ps.find("st", "float");
    unused("trianglemesh", ps);
    const uint32_t mesh = (uint32_t)out_.meshes.size();
    out_.meshes.push_back(std::move(m));
    if (cur_obj_) {
      if (st_.area) warn("area lights are not supported with object instancing; light dropped");
      cur_obj_->push_back(ObjShape{mesh, st_.mat});
      return true;
    }
    add_prim(mesh, st_.mat, Xf());
    if (st_.area) {
      out_.light_prim.push_back((uint32_t)(out_.prim_mesh.size() - 1));
      for (int c = 0; c < 3; ++c) out_.light_radiance.push_back(st_.L[c]);
    }
    return true;
  }

  void add_prim(uint32_t mesh, const Material &m, const Xf &x) {
    out_.prim_mesh.push_back(mesh);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) out_.prim_xform.push_back(x.m.m[r][c]);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) out_.prim_xform.push_back(x.inv.m[r][c]);
    out_.mat_type.push_back(m.type);
    for (int c = 0; c < 3; ++c) {
      out_.kd.push_back(m.kd[c]);
      out_.ks.push_back(m.ks[c]);
    }
    out_.rough.push_back(m.rough);
  }

  bool command(const std::string &cmd, const std::vector<Tok> &t, size_t &i, const std::string &dir, std::string &err,
               int depth) {
    std::vector<float> v;
    std::string name;
    if (cmd == "Identity") {
      st_.ctm = Xf();
    } else if (cmd == "Translate") {
      if (!numbers(t, i, 3, v, err)) return false;
      st_.ctm = st_.ctm * Xf::from(make(1, 0, 0, v[0], 0, 1, 0, v[1], 0, 0, 1, v[2], 0, 0, 0, 1));
    } else if (cmd == "Scale") {
      if (!numbers(t, i, 3, v, err)) return false;
      st_.ctm = st_.ctm * Xf::from(make(v[0], 0, 0, 0, 0, v[1], 0, 0, 0, 0, v[2], 0, 0, 0, 0, 1));
    } else if (cmd == "Rotate") {
      if (!numbers(t, i, 4, v, err)) return false;
      double axis[3] = {v[1], v[2], v[3]};
      st_.ctm = st_.ctm * make_rotation((double)v[0] * M_PI / 180.0, axis);
    } else if (cmd == "LookAt") {
      if (!numbers(t, i, 9, v, err)) return false;
      const double eye[3] = {v[0], v[1], v[2]}, up[3] = {v[6], v[7], v[8]};
      double d[3] = {(double)(v[3] - v[0]), (double)(v[4] - v[1]), (double)(v[5] - v[2])};
      normalize(d);
      st_.ctm = st_.ctm * make_look_at(eye, d, up);
    } else if (cmd == "Transform" || cmd == "ConcatTransform") {
      if (!numbers(t, i, 16, v, err)) return false;
      const Xf x = Xf::from(make(v[0], v[4], v[8], v[12], v[1], v[5], v[9], v[13], v[2], v[6], v[10], v[14], v[3], v[7],
                                 v[11], v[15]));
      st_.ctm = cmd == "Transform" ? x : st_.ctm * x;
    } else if (cmd == "CoordinateSystem") {
      if (!string_arg(t, i, name, err)) return false;
      coord_[name] = st_.ctm;
    } else if (cmd == "CoordSysTransform") {
      if (!string_arg(t, i, name, err)) return false;
      if (coord_.count(name)) st_.ctm = coord_[name];
      else warn("CoordSysTransform: unknown coordinate system \"" + name + "\"");
    } else if (cmd == "AttributeBegin") {
      stack_.push_back(st_);
    } else if (cmd == "AttributeEnd") {
      if (stack_.empty()) {
        warn("AttributeEnd without AttributeBegin");
      } else {
        st_ = stack_.back();
        stack_.pop_back();
      }
    } else if (cmd == "TransformBegin") {
      xf_stack_.push_back(st_.ctm);
    } else if (cmd == "TransformEnd") {
      if (xf_stack_.empty()) {
        warn("TransformEnd without TransformBegin");
      } else {
        st_.ctm = xf_stack_.back();
        xf_stack_.pop_back();
      }
    } else if (cmd == "WorldBegin") {
      st_.ctm = Xf();
      coord_["world"] = st_.ctm;
    } else if (cmd == "WorldEnd") {
      while (!stack_.empty()) {
        warn("missing AttributeEnd");
        st_ = stack_.back();
        stack_.pop_back();
      }
    } else if (cmd == "ReverseOrientation") {
      st_.reverse = !st_.reverse;
    } else if (cmd == "Include") {
      if (!string_arg(t, i, name, err)) return false;
      const std::string p = (!name.empty() && name[0] == '/') ? name : dir + name;
      if (!run(p, err, depth + 1)) return false;
    } else if (cmd == "ObjectBegin") {
      if (!string_arg(t, i, name, err)) return false;
      stack_.push_back(st_);
      if (cur_obj_) {
        warn("ObjectBegin inside an object definition ignored");
      } else {
        objects_[name].clear();
        cur_obj_ = &objects_[name];
      }
    } else if (cmd == "ObjectEnd") {
      if (!cur_obj_) {
        warn("ObjectEnd outside an object definition ignored");
      } else {
        cur_obj_ = nullptr;
        if (!stack_.empty()) {
          st_ = stack_.back();
          stack_.pop_back();
        }
      }
    } else if (cmd == "ObjectInstance") {
      if (!string_arg(t, i, name, err)) return false;
      if (cur_obj_) warn("ObjectInstance inside an object definition ignored");
      else if (!objects_.count(name)) warn("ObjectInstance: unknown object \"" + name + "\"");
      else
        for (const ObjShape &s : objects_[name]) add_prim(s.mesh, s.mat, st_.ctm);
    } else {
      // commands with a name and a parameter list
      if (!string_arg(t, i, name, err)) return false;
      if (cmd == "Texture") {  // "name" "type" "class" params
        std::string a, b;
        if (!string_arg(t, i, a, err) || !string_arg(t, i, b, err)) return false;
      }
      Params ps;
      if (!params(t, i, ps, err)) return false;
      if (cmd == "Shape") return shape(name, ps);
      if (cmd == "Material") {
        st_.mat = material(name, ps);
        unused("Material", ps);
      } else if (cmd == "MakeNamedMaterial") {
        named_[name] = material(ps.one_string("type", "matte"), ps);
        unused("MakeNamedMaterial", ps);
      } else if (cmd == "NamedMaterial") {
        if (named_.count(name)) st_.mat = named_[name];
        else warn("NamedMaterial: unknown material \"" + name + "\"");
      } else if (cmd == "AreaLightSource") {
        if (name == "diffuse" || name == "area") {
          st_.area = true;
          st_.L[0] = st_.L[1] = st_.L[2] = 1.0;
          ps.rgb("L", st_.L);
          double sc[3] = {1.0, 1.0, 1.0};
          ps.rgb("scale", sc);
          for (int c = 0; c < 3; ++c) st_.L[c] *= sc[c];
          ps.one_int("nsamples", 1);
          unused("AreaLightSource", ps);
        } else {
          warn("area light \"" + name + "\" is not supported");
        }
      } else if (cmd == "Film") {
        out_.width = (uint32_t)ps.one_int("xresolution", 640);
        out_.height = (uint32_t)ps.one_int("yresolution", 480);
        film_ = true;
        ps.one_string("filename", "");
        unused("Film", ps);
      } else if (cmd == "Camera") {
        camera(name, ps);
      } else {
        warn("command " + cmd + " \"" + name + "\" ignored");
      }
    }
    return true;
  }

  bool film_ = false;
  Params cam_params_;
  std::string cam_name_;
  Xf cam_to_world_;
  bool cam_seen_ = false;

  void camera(const std::string &name, Params &ps) {
    if (name != "perspective") warn("camera \"" + name + "\" is not supported; perspective used");
    cam_name_ = name;
    cam_params_ = ps;
    cam_to_world_ = st_.ctm.inverts_orientation()
                        ? (Xf::from(make(-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)) * st_.ctm).inverted()
                        : st_.ctm.inverted();
    coord_["camera"] = cam_to_world_;
    cam_seen_ = true;
  }

 public:
  // The camera after the whole file is read (Film may follow Camera).
  void finish() {
    if (!cam_seen_) return;
    Params &ps = cam_params_;
    float fov = ps.one_float("fov", 90.0f);
    const float halffov = ps.one_float("halffov", -1.0f);
    if (halffov > 0.0f) fov = 2.0f * halffov;
    const double frame = (double)ps.one_float("frameaspectratio", out_.width / (float)out_.height);
    const double half_tan = std::tan((double)fov * M_PI / 180.0 / 2.0);
    const double x_angle = std::atan(half_tan * std::max(1.0, frame)) * 2.0;
    if (ps.one_float("lensradius", 0.0f) != 0.0f) warn("camera: lensradius > 0 (depth of field) is not supported");
    ps.one_float("focaldistance", 0.0f);
    ps.find("screenwindow", "float");
    unused("Camera", ps);
    skw_camera &c = out_.cam;
    for (int r = 0; r < 3; ++r) {
      c.origin[r] = cam_to_world_.m.m[r][3];
      for (int k = 0; k < 3; ++k) c.rot[3 * r + k] = cam_to_world_.m.m[r][k];
    }
    c.x_tan = std::tan(x_angle / 2.0);
    c.y_tan = c.x_tan * ((double)out_.height) / out_.width;
    out_.has_camera = true;
  }
};

}  // namespace

extern "C" {

skw_status skw_pbrt_load(const char *path, skw_pbrt_scene **out) {
  if (!path || !out) return fail(SKW_INVALID_ARG, "path and out must not be NULL");
  *out = nullptr;
  std::unique_ptr<skw_pbrt_scene> s(new skw_pbrt_scene());
  std::memset(&s->cam, 0, sizeof(s->cam));
  Importer im(*s);
  std::string err;
  if (!im.run(path, err)) return fail(SKW_INVALID_ARG, std::string("pbrt import: ") + err);
  im.finish();
  // flatten the meshes for the scene description
  s->vo.push_back(0);
  s->to.push_back(0);
  bool any_normals = false;
  for (const Mesh &m : s->meshes) any_normals = any_normals || !m.normals.empty();
  for (const Mesh &m : s->meshes) {
    s->verts.insert(s->verts.end(), m.verts.begin(), m.verts.end());
    if (any_normals) {
      if (m.normals.empty()) s->normals.insert(s->normals.end(), m.verts.size(), 0.0f);
      else s->normals.insert(s->normals.end(), m.normals.begin(), m.normals.end());
    }
    s->tris.insert(s->tris.end(), m.tris.begin(), m.tris.end());
    s->flags.push_back(m.flags);
    s->vo.push_back(s->verts.size() / 3);
    s->to.push_back(s->tris.size() / 3);
  }
  *out = s.release();
  return SKW_OK;
}

skw_status skw_pbrt_descs(const skw_pbrt_scene *s, skw_scene_desc *scene, skw_dl_desc *dl, skw_camera *cam,
                          uint32_t *width, uint32_t *height) {
  if (!s) return fail(SKW_INVALID_ARG, "scene must not be NULL");
  if (scene) {
    std::memset(scene, 0, sizeof(*scene));
    scene->n_meshes = (uint32_t)s->meshes.size();
    scene->mesh_vert_offset = s->vo.data();
    scene->verts = s->verts.data();
    scene->normals = s->normals.empty() ? nullptr : s->normals.data();
    scene->mesh_tri_offset = s->to.data();
    scene->tris = s->tris.data();
    scene->mesh_flags = s->flags.data();
    scene->n_prims = (uint32_t)s->prim_mesh.size();
    scene->prim_mesh = s->prim_mesh.data();
    scene->prim_xform = s->prim_xform.data();
  }
  if (dl) {
    std::memset(dl, 0, sizeof(*dl));
    dl->mat_type = s->mat_type.data();
    dl->kd = s->kd.data();
    dl->ks = s->ks.data();
    dl->roughness = s->rough.data();
    dl->n_lights = (uint32_t)s->light_prim.size();
    dl->light_prim = s->light_prim.data();
    dl->light_radiance = s->light_radiance.data();
  }
  if (cam) *cam = s->cam;
  if (width) *width = s->width;
  if (height) *height = s->height;
  if (cam && !s->has_camera) return fail(SKW_INVALID_ARG, "the file defines no camera");
  return SKW_OK;
}

skw_status skw_pbrt_warnings(const skw_pbrt_scene *s, uint32_t *count) {
  if (!s || !count) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  *count = (uint32_t)s->warnings.size();
  return SKW_OK;
}

skw_status skw_pbrt_warning(const skw_pbrt_scene *s, uint32_t i, const char **msg) {
  if (!s || !msg) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  if (i >= s->warnings.size()) return fail(SKW_INVALID_ARG, "warning index out of range");
  *msg = s->warnings[i].c_str();
  return SKW_OK;
}

skw_status skw_pbrt_destroy(skw_pbrt_scene *s) {
  delete s;
  return SKW_OK;
}

}  // extern "C"
