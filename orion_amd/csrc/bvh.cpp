// Binned-SAH BVH build on the host (replaces Embree's rtcCommit,
// core/kernel.cpp:108).  Subtrees above 64k triangles build in parallel.
#include "bvh.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <string>
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

int leaf_code(const TNode &l) { return ~(l.first * 8 + (l.count - 1)); }

// Grid step of one axis: the smallest power of two 2^e with
// fmaf(255, 2^e, lo) >= hi, as a biased exponent byte.
uint32_t grid_exponent(float lo, float hi) {
  double ext = (double)hi - (double)lo;
  int e = ext > 0 ? (int)std::ceil(std::log2(ext / 255.0)) : -126;
  e = std::max(-126, std::min(127, e));
  for (;;) {
    float step;
    uint32_t bits = (uint32_t)(e + 127) << 23;
    std::memcpy(&step, &bits, 4);
    if (std::fmaf(255.0f, step, lo) >= hi || e >= 127) break;
    ++e;
  }
  return (uint32_t)(e + 127);
}

float step_of(uint32_t ebyte) {
  uint32_t bits = ebyte << 23;
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// Quantizes [clo, chi] on one axis to bytes whose decoded planes enclose it.
void quantize(float clo, float chi, float org, float step, uint8_t *qlo, uint8_t *qhi) {
  int lo = (int)std::floor(((double)clo - org) / step);
  lo = std::max(0, std::min(255, lo));
  while (lo > 0 && std::fmaf((float)lo, step, org) > clo) --lo;
  int hi = (int)std::ceil(((double)chi - org) / step);
  hi = std::max(0, std::min(255, hi));
  while (hi < 255 && std::fmaf((float)hi, step, org) < chi) ++hi;
  *qlo = (uint8_t)lo;
  *qhi = (uint8_t)hi;
}

// Collapses the binary tree into 4-wide nodes (a node's slots take the
// children of its largest-area inner child until there are four) and writes
// them breadth first, siblings contiguous.
void emit_wide(const std::vector<TNode> &t, Bvh *out) {
  out->nodes.clear();
  out->depth = 0;
  out->stack_bound = 1;
  BvhNode root{};
  if (t.empty()) {
    for (int k = 0; k < 4; ++k) root.child[k] = kLeafEmpty;
    for (int k = 0; k < 4; ++k) { root.q[k] = 255; root.q[4 + k] = 0; }
    out->nodes.push_back(root);
    return;
  }
  out->nodes.push_back(root);
  struct Item { int tnode, slot, depth; };
  std::vector<Item> queue;
  queue.push_back({0, 0, 0});
  for (size_t qi = 0; qi < queue.size(); ++qi) {
    Item it = queue[qi];
    out->depth = std::max(out->depth, it.depth);
    int kids[4], nk = 0;
    if (t[it.tnode].count > 0) {
      kids[nk++] = it.tnode;  // a lone leaf at the root
    } else {
      kids[nk++] = t[it.tnode].kid[0];
      kids[nk++] = t[it.tnode].kid[1];
      while (nk < 4) {
        int best = -1;
        float ba = -1;
        for (int k = 0; k < nk; ++k)
          if (t[kids[k]].count == 0 && t[kids[k]].box.area() > ba) { ba = t[kids[k]].box.area(); best = k; }
        if (best < 0) break;
        int open = kids[best];
        kids[best] = t[open].kid[0];
        kids[nk++] = t[open].kid[1];
      }
    }
    Box nb;
    for (int k = 0; k < nk; ++k) nb.grow(t[kids[k]].box);
    BvhNode g{};
    uint32_t eb[3];
    float step[3];
    for (int a = 0; a < 3; ++a) {
      g.org[a] = nb.lo[a];
      eb[a] = grid_exponent(nb.lo[a], nb.hi[a]);
      step[a] = step_of(eb[a]);
    }
    g.exps = eb[0] | (eb[1] << 8) | (eb[2] << 16);
    for (int k = 0; k < 4; ++k) {
      if (k >= nk) {
        g.child[k] = kLeafEmpty;
        for (int a = 0; a < 3; ++a) { g.q[8 * a + k] = 255; g.q[8 * a + 4 + k] = 0; }
        continue;
      }
      const TNode &c = t[kids[k]];
      for (int a = 0; a < 3; ++a) quantize(c.box.lo[a], c.box.hi[a], g.org[a], step[a], &g.q[8 * a + k], &g.q[8 * a + 4 + k]);
      if (c.count > 0) {
        g.child[k] = leaf_code(c);
      } else {
        g.child[k] = (int)out->nodes.size();
        out->nodes.push_back(BvhNode{});
        queue.push_back({kids[k], g.child[k], it.depth + 1});
      }
    }
    out->nodes[it.slot] = g;
  }
  // a depth-first walk keeps at most three siblings per level waiting
  // (levels 1 .. depth + 1)
  out->stack_bound = 3 * (out->depth + 1);
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
  // triangles in leaf order; leaves are encoded in their parent's child slot
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
  emit_wide(t, out);
  out->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

std::string check_bvh(const Bvh &b, const float *tri_pn, int n) {
  if (b.nodes.empty()) return "no root node";
  if ((int)b.tris.size() != n || (int)b.prim.size() != n) return "triangle count mismatch";
  std::vector<int> seen(n, 0);
  struct Item { int node, depth; };
  std::vector<Item> st{{0, 0}};
  int max_depth = 0;
  std::vector<char> visited(b.nodes.size(), 0);
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    if (it.node < 0 || it.node >= (int)b.nodes.size()) return "child index out of range";
    if (visited[it.node]++) return "node reached twice";
    max_depth = std::max(max_depth, it.depth);
    const BvhNode &g = b.nodes[it.node];
    float step[3];
    for (int a = 0; a < 3; ++a) step[a] = step_of((g.exps >> (8 * a)) & 0xffu);
    for (int k = 0; k < 4; ++k) {
      int c = g.child[k];
      if (c == kLeafEmpty) continue;
      float lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::fmaf((float)g.q[8 * a + k], step[a], g.org[a]);
        hi[a] = std::fmaf((float)g.q[8 * a + 4 + k], step[a], g.org[a]);
      }
      // every triangle below this slot lies inside the decoded box
      std::vector<int> sub{c};
      while (!sub.empty()) {
        int x = sub.back();
        sub.pop_back();
        if (x >= 0) {
          if (x >= (int)b.nodes.size()) return "child index out of range";
          for (int kk = 0; kk < 4; ++kk)
            if (b.nodes[x].child[kk] != kLeafEmpty) sub.push_back(b.nodes[x].child[kk]);
          continue;
        }
        int code = ~x, first = code >> 3, cnt = (code & 7) + 1;
        if (first < 0 || first + cnt > n) return "leaf range out of bounds";
        for (int s2 = first; s2 < first + cnt; ++s2) {
          const float *q = tri_pn + 18 * (size_t)b.prim[s2];
          for (int v = 0; v < 3; ++v)
            for (int a = 0; a < 3; ++a)
              if (!(q[3 * v + a] >= lo[a] && q[3 * v + a] <= hi[a])) return "decoded box misses a triangle";
        }
      }
      if (c >= 0) {
        st.push_back({c, it.depth + 1});
      } else {
        int code = ~c, first = code >> 3, cnt = (code & 7) + 1;
        for (int s2 = first; s2 < first + cnt; ++s2) seen[s2]++;
      }
    }
  }
  for (int i = 0; i < n; ++i)
    if (seen[i] != 1) return "triangle slot not referenced exactly once";
  for (size_t i = 0; i < visited.size(); ++i)
    if (!visited[i]) return "unreachable node";
  if (max_depth != b.depth) return "depth mismatch";
  if (b.stack_bound != 3 * (b.depth + 1)) return "stack bound mismatch";
  return "";
}

}  // namespace orion_amd
