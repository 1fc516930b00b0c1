// Host-built SAH BVH for the device traversal (replaces Embree's scene build,
// core/kernel.cpp:63-110).
//
// Node format: 4-wide, 64 B, child boxes quantized to 8 bits per plane
// relative to the node's own grid:
//   float org[3]          grid origin
//   uint32 exps           byte a = biased exponent e_a of the grid step
//                         2^(e_a - 127) on axis a (a power of two)
//   int child[4]          >= 0 inner node index; kLeafEmpty = no child;
//                         other negatives: leaf ~(first * 8 + count - 1)
//   uint8 q[24]           lo_x[4] hi_x[4] lo_y[4] hi_y[4] lo_z[4] hi_z[4]
//   int pad[2]
// A child plane decodes as fmaf(q, step, org) in float (one rounding).  The
// builder picks each q so the decoded box contains the child's padded box,
// evaluating that same fmaf, so decoding is conservative bit for bit on
// the device.  Triangles in leaf order, 48 B each: (p0.xyz, prim id bits)
// (e1.xyz, 0) (e2.xyz, 0) with e1 = p1 - p0, e2 = p2 - p0 evaluated in float.
// Boxes are padded outward so the box test never culls a triangle the
// triangle test would accept.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace orion_amd {

constexpr int kBvhMaxDepth = 40;     // binary build depth bound
constexpr int kBvhMaxLeaf = 8;
constexpr int kLeafEmpty = (int)0x80000000;  // empty child slot / end of traversal

struct BvhNode {
  float org[3];
  uint32_t exps;
  int child[4];
  uint8_t q[24];
  int pad[2];
};
static_assert(sizeof(BvhNode) == 64, "node size");

struct BvhTri {
  float v[12];
};

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<BvhTri> tris;   // leaf order
  std::vector<int> prim;      // leaf slot -> primitive id
  int depth = 0;        // 4-wide levels below the root
  int stack_bound = 1;  // traversal stack entries that can be live at once
  double build_ms = 0;
};

// tri_pn: n x 18 floats (p0 p1 p2 n0 n1 n2)
void build_bvh(const float *tri_pn, int n, Bvh *out);
// Structural check: every node reachable once, every triangle slot in one
// leaf, every decoded child box contains its subtree's vertices, depth and
// stack bound consistent.  Returns "" when the tree is well formed.
std::string check_bvh(const Bvh &b, const float *tri_pn, int n);

}  // namespace orion_amd
