// Binned-SAH BVH build on the host (replaces Embree's rtcCommit,
// core/kernel.cpp:108).  Subtrees above 64k triangles build in parallel.
#include "bvh.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <vector>

namespace orion_amd {
namespace {

constexpr int kBins = 32;
constexpr float kInfF = std::numeric_limits<float>::infinity();

struct Box {
  float lo[3] = {kInfF, kInfF, kInfF}, hi[3] = {-kInfF, -kInfF, -kInfF};
  void grow(const Box &b) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  void grow(const float *p) {
    for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (!(dx >= 0)) return 0;
    return 2 * (dx * dy + dy * dz + dz * dx);
  }
};

struct TNode {
  Box box;
  int kid[2] = {-1, -1};
  int first = 0, count = 0;  // count > 0: leaf
};

struct Ctx {
  std::vector<Box> tb;        // padded triangle boxes
  std::vector<float> cen;     // centroids (3 per tri)
  std::vector<int> order;
};

int ceil_log2(int x) {
  int r = 0;
  while ((1 << r) < x) ++r;
  return r;
}

// Builds the subtree over order[first, first+count) and returns it as a node
// list whose element 0 is the subtree root (children indices are local).
std::vector<TNode> build(Ctx &c, int first, int count, int depth) {
  std::vector<TNode> out(1);
  TNode &root = out[0];
  Box cb;
  for (int i = first; i < first + count; ++i) {
    int t = c.order[i];
    root.box.grow(c.tb[t]);
    cb.grow(&c.cen[3 * t]);
  }
  int need = ceil_log2((count + kBvhMaxLeaf - 1) / kBvhMaxLeaf);
  bool forced = depth + need >= kBvhMaxDepth;
  if (count <= 2 || (count <= kBvhMaxLeaf && depth >= kBvhMaxDepth)) {
    out[0].first = first; out[0].count = count;
    return out;
  }
  int axis = -1, split_bin = -1, mid = -1;
  float best = kInfF;
  if (!forced) {
    for (int a = 0; a < 3; ++a) {
      float ext = cb.hi[a] - cb.lo[a];
      if (!(ext > 0)) continue;
      float scale = kBins / ext;
      int cnt[kBins] = {0};
      Box bb[kBins];
      for (int i = first; i < first + count; ++i) {
        int t = c.order[i];
        int k = std::min(kBins - 1, std::max(0, (int)((c.cen[3 * t + a] - cb.lo[a]) * scale)));
        cnt[k]++;
        bb[k].grow(c.tb[t]);
      }
      float lA[kBins];
      int lC[kBins];
      Box acc;
      int n = 0;
      for (int k = 0; k < kBins; ++k) { acc.grow(bb[k]); n += cnt[k]; lA[k] = acc.area(); lC[k] = n; }
      acc = Box();
      for (int k = kBins - 1; k > 0; --k) {
        acc.grow(bb[k]);
        int nl = lC[k - 1], nr = count - nl;
        if (nl == 0 || nr == 0) continue;
        float cost = lA[k - 1] * nl + acc.area() * nr;
        if (cost < best) { best = cost; axis = a; split_bin = k - 1; }
      }
    }
    float parent = root.box.area();
    float leaf_cost = (float)count;
    float split_cost = parent > 0 ? 0.3f + best / parent : kInfF;
    if (count <= kBvhMaxLeaf && (axis < 0 || leaf_cost <= split_cost)) {
      out[0].first = first; out[0].count = count;
      return out;
    }
  }
  if (axis >= 0) {
    float ext = cb.hi[axis] - cb.lo[axis];
    float scale = kBins / ext;
    auto it = std::partition(c.order.begin() + first, c.order.begin() + first + count, [&](int t) {
      int k = std::min(kBins - 1, std::max(0, (int)((c.cen[3 * t + axis] - cb.lo[axis]) * scale)));
      return k <= split_bin;
    });
    mid = (int)(it - c.order.begin());
  }
  if (axis < 0 || mid <= first || mid >= first + count) {
    // object median on the widest centroid axis (also the depth-bound fallback)
    int a = 0;
    for (int q = 1; q < 3; ++q)
      if (cb.hi[q] - cb.lo[q] > cb.hi[a] - cb.lo[a]) a = q;
    mid = first + count / 2;
    std::nth_element(c.order.begin() + first, c.order.begin() + mid, c.order.begin() + first + count,
                     [&](int x, int y) { return c.cen[3 * x + a] < c.cen[3 * y + a]; });
  }
  std::vector<TNode> L, R;
  if (count > 65536 && depth < 8) {
    auto fl = std::async(std::launch::async, [&] { return build(c, first, mid - first, depth + 1); });
    R = build(c, mid, first + count - mid, depth + 1);
    L = fl.get();
  } else {
    L = build(c, first, mid - first, depth + 1);
    R = build(c, mid, first + count - mid, depth + 1);
  }
  auto append = [&](std::vector<TNode> &sub) {
    int off = (int)out.size();
    for (auto &n : sub) {
      for (int k = 0; k < 2; ++k)
        if (n.kid[k] >= 0) n.kid[k] += off;
      out.push_back(n);
    }
    return off;
  };
  int l = append(L);
  int r = append(R);
  out[0].kid[0] = l;
  out[0].kid[1] = r;
  return out;
}

}  // namespace

void build_bvh(const float *tri_pn, int n, Bvh *out) {
  auto t0 = std::chrono::steady_clock::now();
  Ctx c;
  c.tb.resize(n);
  c.cen.resize(3 * (size_t)n);
  c.order.resize(n);
  for (int i = 0; i < n; ++i) {
    const float *q = tri_pn + 18 * (size_t)i;
    Box b;
    b.grow(q); b.grow(q + 3); b.grow(q + 6);
    for (int a = 0; a < 3; ++a) {
      // outward padding: the triangle test's rounding may accept a hit just
      // outside the exact box; the box test must never cull it
      float pad = (std::max(std::abs(b.lo[a]), std::abs(b.hi[a])) + (b.hi[a] - b.lo[a])) * 2e-5f + 1e-20f;
      b.lo[a] -= pad;
      b.hi[a] += pad;
      c.cen[3 * (size_t)i + a] = 0.5f * (b.lo[a] + b.hi[a]);
    }
    c.tb[i] = b;
    c.order[i] = i;
  }
  std::vector<TNode> t;
  if (n > 0) t = build(c, 0, n, 0);
  // flatten: inner nodes get GPU node slots in depth-first order; leaves are
  // encoded in the parent's child field
  out->nodes.clear();
  out->tris.resize(n);
  out->prim = c.order;
  for (int s = 0; s < n; ++s) {
    int p = c.order[s];
    const float *q = tri_pn + 18 * (size_t)p;
    BvhTri &bt = out->tris[s];
    int pid = p;
    std::memcpy(&bt.v[3], &pid, 4);
    for (int a = 0; a < 3; ++a) {
      bt.v[a] = q[a];
      bt.v[4 + a] = q[3 + a] - q[a];
      bt.v[8 + a] = q[6 + a] - q[a];
    }
    bt.v[7] = 0; bt.v[11] = 0;
  }
  auto leaf_code = [](const TNode &l) { return ~(l.first * 8 + (l.count - 1)); };
  auto empty_box = [](float *dst) { dst[0] = kInfF; dst[1] = -kInfF; dst[2] = kInfF; dst[3] = -kInfF; };
  int max_depth = 0;
  if (n == 0) {
    BvhNode e;
    empty_box(e.a); empty_box(e.b);
    e.c[0] = kInfF; e.c[1] = -kInfF; e.c[2] = kInfF; e.c[3] = -kInfF;
    e.d[0] = e.d[1] = ~0; e.d[2] = e.d[3] = 0;
    out->nodes.push_back(e);
  } else if (t[0].count > 0) {
    // single leaf: a root whose second child is empty
    BvhNode e;
    const Box &b = t[0].box;
    e.a[0] = b.lo[0]; e.a[1] = b.hi[0]; e.a[2] = b.lo[1]; e.a[3] = b.hi[1];
    empty_box(e.b);
    e.c[0] = b.lo[2]; e.c[1] = b.hi[2]; e.c[2] = kInfF; e.c[3] = -kInfF;
    e.d[0] = leaf_code(t[0]); e.d[1] = leaf_code(t[0]); e.d[2] = e.d[3] = 0;
    out->nodes.push_back(e);
  } else {
    std::vector<int> slot(t.size(), -1);
    std::vector<std::pair<int, int>> stack;  // (tnode, depth)
    stack.push_back({0, 0});
    // first pass: assign slots in DFS pre-order (child 0 first)
    int next = 0;
    std::vector<int> dfs;
    while (!stack.empty()) {
      auto [i, d] = stack.back();
      stack.pop_back();
      if (t[i].count > 0) continue;
      slot[i] = next++;
      dfs.push_back(i);
      max_depth = std::max(max_depth, d);
      stack.push_back({t[i].kid[1], d + 1});
      stack.push_back({t[i].kid[0], d + 1});
    }
    out->nodes.resize(next);
    for (int i : dfs) {
      BvhNode &g = out->nodes[slot[i]];
      const TNode &c0 = t[t[i].kid[0]], &c1 = t[t[i].kid[1]];
      g.a[0] = c0.box.lo[0]; g.a[1] = c0.box.hi[0]; g.a[2] = c0.box.lo[1]; g.a[3] = c0.box.hi[1];
      g.b[0] = c1.box.lo[0]; g.b[1] = c1.box.hi[0]; g.b[2] = c1.box.lo[1]; g.b[3] = c1.box.hi[1];
      g.c[0] = c0.box.lo[2]; g.c[1] = c0.box.hi[2]; g.c[2] = c1.box.lo[2]; g.c[3] = c1.box.hi[2];
      g.d[0] = c0.count > 0 ? leaf_code(c0) : slot[t[i].kid[0]];
      g.d[1] = c1.count > 0 ? leaf_code(c1) : slot[t[i].kid[1]];
      g.d[2] = g.d[3] = 0;
    }
  }
  out->depth = max_depth;
  out->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace orion_amd
