// Device data layouts shared by the host builders and the gfx950 kernels.
//
// Triangle tree (TriangleAccelerator): nodes are stored so that the children of a node occupy one
// contiguous block in left, middle, right order, which lets a traversal stack entry name "the next
// sibling and how many are left" in one 32-bit word. Triangles are stored in the order the tree's
// in-place partition leaves them (TriangleAccelerator.cpp:644-701), so a leaf is one contiguous run.
#pragma once
#include <cstdint>

namespace skw {

// 32 bytes. b: bits 0-1 = 0 for a leaf, split axis + 1 for an internal node.
//   internal: a = index of the first child, bits 2-4 = child mask (left, middle, right).
//   leaf:     a = first triangle position, bit 2 = has instances, bits 3-31 = triangle count.
struct alignas(16) DevNode {
  float bmin[3];
  float bmax[3];
  uint32_t a;
  uint32_t b;
};
static_assert(sizeof(DevNode) == 32, "node is 32 bytes");

// 48 bytes: vertex floats in the tree's order plus the ids reported for a hit.
struct alignas(16) DevTri {
  float v0[3];
  float v1[3];
  float v2[3];
  uint32_t prim;  // 0xFFFFFFFF for triangles of an instanced mesh
  uint32_t tri;   // triangle index inside its mesh
  uint32_t pad;
};
static_assert(sizeof(DevTri) == 48, "triangle is 48 bytes");

// One instanced primitive: world-to-mesh matrix (row-major 4x4), root node of the mesh's subtree
// and the primitive id.
struct alignas(16) DevInst {
  double w2m[16];
  uint32_t root;
  uint32_t prim;
  uint32_t pad[2];
};

// Per primitive data the hit-attribute kernels need (mesh-space differential geometry).
struct alignas(16) DevPrim {
  double m2w[16];
  double w2m[16];
  uint32_t mesh;
  uint32_t identity;
  uint32_t pad[2];
};

struct DevMesh {
  uint64_t vert_off;
  uint64_t tri_off;
  uint32_t has_normals;
  uint32_t invert_normals;
};

// Photon kd-tree: one 16-byte record per node = the node's photon position plus the packed word
// of KDTree::Node (KDTree.h:187-246): bits 0-1 split axis (3 = leaf), bit 2 has-left-child,
// bits 3-31 right child index ((1<<29)-1 = none). The split coordinate is the node photon's own
// coordinate on the split axis (KDTree.h:456), so it needs no separate field.
struct alignas(16) DevKdNode {
  float p[3];
  uint32_t word;
};
static_assert(sizeof(DevKdNode) == 16, "kd node is 16 bytes");

// Photon payload in tree order: weight RGB, incident direction and normal ids, original index.
struct alignas(16) DevPhoton {
  float w[3];
  uint32_t codes;  // dir | normal << 16
};

constexpr uint32_t kNoChild = (1u << 29) - 1;
constexpr uint32_t kMiss = 0xFFFFFFFFu;

}  // namespace skw
