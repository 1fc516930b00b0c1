// Host-built SAH BVH for the device traversal (replaces Embree's scene build,
// core/kernel.cpp:63-110).
//
// Node format (64 B, both children's boxes stored in the parent so one node
// fetch tests two boxes):
//   float4 a = c0.lo.x c0.hi.x c0.lo.y c0.hi.y
//   float4 b = c1.lo.x c1.hi.x c1.lo.y c1.hi.y
//   float4 c = c0.lo.z c0.hi.z c1.lo.z c1.hi.z
//   int4   d = child0 child1 - -     child >= 0: node index;
//                                    child <  0: leaf ~(first * 8 + count - 1)
// Triangles in leaf order, 48 B each: (p0.xyz, prim id bits) (e1.xyz, 0)
// (e2.xyz, 0) with e1 = p1 - p0, e2 = p2 - p0 evaluated in float.
// Boxes are padded outward so the box test never culls a triangle the
// triangle test would accept; depth is bounded by kMaxDepth so a fixed
// traversal stack of kMaxDepth entries cannot overflow.
#pragma once
#include <cstdint>
#include <vector>

namespace orion_amd {

constexpr int kBvhMaxDepth = 40;     // levels below the root (stack entries needed)
constexpr int kBvhMaxLeaf = 8;

struct BvhNode {
  float a[4], b[4], c[4];
  int d[4];
};
static_assert(sizeof(BvhNode) == 64, "node size");

struct BvhTri {
  float v[12];
};

struct Bvh {
  std::vector<BvhNode> nodes;
  std::vector<BvhTri> tris;   // leaf order
  std::vector<int> prim;      // leaf slot -> primitive id
  int depth = 0;
  double build_ms = 0;
};

// tri_pn: n x 18 floats (p0 p1 p2 n0 n1 n2)
void build_bvh(const float *tri_pn, int n, Bvh *out);

}  // namespace orion_amd
