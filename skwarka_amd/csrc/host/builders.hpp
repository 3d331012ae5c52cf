// Host-side builders: the ternary triangle tree and the photon kd-tree, producing the flattened
// device layouts of skw_layout.hpp.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../skw_layout.hpp"
#include "../../../include/skwarka_rt.h"

namespace skw {

struct HostMesh {
  std::vector<float> verts;      // 3 per vertex
  std::vector<float> normals;    // 3 per vertex (normalised as float), may be empty
  std::vector<uint32_t> tris;    // 3 per triangle after cleaning
  bool invert_normals = false;
};

struct HostPrim {
  double m2w[16];
  double w2m[16];
  uint32_t mesh;
  bool identity;
};

struct TreeLayout {
  std::vector<DevNode> nodes;
  std::vector<DevTri> tris;
  std::vector<DevInst> insts;
  std::vector<uint32_t> inst_range;  // 2 per node (begin, end) when the scene has instances
  uint32_t root = 0;
  uint32_t max_depth = 0;   // deepest node (root = 0)
  uint32_t stack_need = 0;  // stack entries a traversal needs (outer + instanced subtree)
  uint64_t leaves = 0;
  float bounds[6] = {0, 0, 0, 0, 0, 0};
  // Every child's box lies inside its parent's box and no box coordinate is NaN. True for every
  // tree the builders make (a node's box is the fold of its triangles' boxes, which start at
  // (+inf, -inf) and never take a NaN); checked for trees read from files.
  bool nested = true;
  std::vector<HostMesh> meshes;
  std::vector<HostPrim> prims;
};

// Reads and cleans the description, builds the tree. Returns false with a message on bad input.
bool build_triangle_tree(const skw_scene_desc &desc, TreeLayout &out, std::string &err);

// The tree builder's input without the build, for scenes without instanced meshes: meshes and
// primitives read and cleaned; goff: per mesh the first triangle of the builder's order (meshes
// used by one primitive, in mesh order; n_meshes + 1 entries); user: the primitive using each
// mesh. Sets instanced (offsets left empty) when a mesh is used by several primitives.
bool read_plain_scene(const skw_scene_desc &desc, TreeLayout &out, std::vector<uint64_t> &goff,
                      std::vector<uint32_t> &user, bool &instanced, std::string &err);

// Depth-first pre-order export (children left, middle, right) in the form of skw_scene_export_tree.
uint64_t export_tree(const TreeLayout &t, float *bbox, uint64_t *meta, uint32_t *tri_prim, uint32_t *tri_tri);

struct KdLayout {
  std::vector<DevKdNode> nodes;
  std::vector<DevPhoton> photons;
  std::vector<uint32_t> perm;  // tree position -> caller's photon index
  std::vector<float> psplit;   // tree position -> the parent's split coordinate (0 for the root)
  uint32_t max_depth = 0;
};

void build_photon_tree(uint64_t n, const float *pos, const float *weight, const uint16_t *dir,
                       const uint16_t *normal, KdLayout &out);

// CompressedDirection decode table (Source/RayLibs/Math/CompressedDirection.cpp:20-44).
const double *cdir_table();

}  // namespace skw
