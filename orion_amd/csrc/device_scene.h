// Device scene layout in HBM and the BVH traversal (replaces rtcIntersect,
// core/kernel.cpp:111-174).
//
// Closest-hit contract (shared with the CPU side): Moller-Trumbore in float
// without contraction on (p0, e1 = p1 - p0, e2 = p2 - p0); a triangle is
// accepted when tmin < t <= tmax; among accepted triangles the hit is the
// lexicographic minimum of (t, primitive id), so the answer does not depend
// on traversal order.  `tie` flags rays where a second triangle was accepted
// at exactly the winning t (shared edges / vertices).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bvh.h"
#include "device_math.h"

namespace orion_amd {

constexpr int kStackEntries = kBvhMaxDepth + 2;  // worst case; kernels use the scene's depth + 2

struct DScene {
  const float4 *nodes;   // 4 x float4 per node (bvh.h)
  const float4 *tris;    // 3 x float4 per leaf slot
  const float4 *shade;   // 5 x float4 per primitive: p0 p1 p2 n0 n1 n2 (mat, light bits)
  const float *mats;     // 16 floats per material
  const float *lights;   // 32 floats per light
  const float *light_cdf;   // n_lights + 1
  const float *light_func;  // n_lights
  const float *light_area;  // n_lights (area lights; 0 otherwise)
  const uint32_t *sobol;    // 1024 x 32 generator columns
  float light_func_int;
  int n_lights;
  int n_prims;
  int stack_entries;  // traversal stack depth this scene needs (BVH depth + 2)
  float cam[16];
};

namespace dev {

// Robust slab test (Ize 2013 / pbrt-v3): the exit distance is scaled by
// 1 + 2 gamma(3) so rounding never culls a box the exact test would enter.
// The scale is applied once to min(t_far) -- identical to scaling each slab
// first, as multiplication by a positive constant is monotonic.
__device__ __forceinline__ bool box_hit(float lox, float hix, float loy, float hiy, float loz, float hiz,
                                        V3 o, V3 inv, float tmin, float tmax, float *tnear) {
  float ax = (lox - o.x) * inv.x, bx = (hix - o.x) * inv.x;
  float ay = (loy - o.y) * inv.y, by = (hiy - o.y) * inv.y;
  float az = (loz - o.z) * inv.z, bz = (hiz - o.z) * inv.z;
  float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
  float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz)) * 1.00000035762786865234375f;
  *tnear = t0;
  return t0 <= fminf(t1, tmax);
}

// Triangle test: identical operation sequence to the CPU restatement.
__device__ __forceinline__ bool tri_test(float4 a, float4 b, float4 c, V3 o, V3 d, float tmin, float tmax,
                                         float *t, float *u, float *v) {
  V3 p0 = mk(a.x, a.y, a.z), e1 = mk(b.x, b.y, b.z), e2 = mk(c.x, c.y, c.z);
  V3 pv = cross(d, e2);
  float det = dot(e1, pv);
  if (det == 0.0f || det != det) return false;
  float inv = 1.0f / det;
  V3 tv = sub(o, p0);
  float uu = dot(tv, pv) * inv;
  if (!(uu >= 0.0f && uu <= 1.0f)) return false;
  V3 qv = cross(tv, e1);
  float vv = dot(d, qv) * inv;
  if (!(vv >= 0.0f && uu + vv <= 1.0f)) return false;
  float tt = dot(e2, qv) * inv;
  if (!(tt > tmin && tt <= tmax)) return false;
  *t = tt; *u = uu; *v = vv;
  return true;
}

struct HitRec {
  float t, u, v;
  int prim;
  int tie;
};

// Stack: `stk` points at this lane's first entry, consecutive entries are
// `stride` ints apart (LDS column per lane: conflict-free across a wave).
template <bool ANY>
__device__ inline bool traverse(const DScene &sc, V3 o, V3 d, float tmin, float tmax, HitRec *h, int *stk,
                                int stride, uint32_t *node_visits, uint32_t *tri_tests) {
  // |d| components below 1e-35 are clamped so (lo - o) * inv is never 0 * inf
  // (a NaN would drop a slab); the padded boxes absorb the change
  auto safe = [](float x) { return fabsf(x) > 1e-35f ? x : copysignf(1e-35f, x); };
  V3 inv = mk(1.0f / safe(d.x), 1.0f / safe(d.y), 1.0f / safe(d.z));
  bool found = false;
  float best = tmax;
  h->prim = -1;
  h->tie = 0;
  int sp = 0;
  int node = 0;
  uint32_t nv = 0, nt = 0;
  for (;;) {
    if (node >= 0) {
      // inner node: test both children
      const float4 *nd = sc.nodes + 4 * (size_t)node;
      float4 A = nd[0], B = nd[1], C = nd[2];
      int4 D = *reinterpret_cast<const int4 *>(nd + 3);
      ++nv;
      float t0, t1;
      // an empty child slot has lo.x > hi.x and is never entered
      bool h0 = A.x <= A.y && box_hit(A.x, A.y, A.z, A.w, C.x, C.y, o, inv, tmin, best, &t0);
      bool h1 = B.x <= B.y && box_hit(B.x, B.y, B.z, B.w, C.z, C.w, o, inv, tmin, best, &t1);
      if (h0 && h1) {
        int nearc = D.x, farc = D.y;
        if (t1 < t0) { nearc = D.y; farc = D.x; }
        stk[stride * sp++] = farc;
        node = nearc;
        continue;
      } else if (h0) {
        node = D.x;
        continue;
      } else if (h1) {
        node = D.y;
        continue;
      }
    } else {
      // leaf
      int code = ~node;
      int first = code >> 3, cnt = (code & 7) + 1;
      for (int k = first; k < first + cnt; ++k) {
        const float4 *tr = sc.tris + 3 * (size_t)k;
        float4 a = tr[0], b = tr[1], c = tr[2];
        ++nt;
        float t, u, v;
        if (!tri_test(a, b, c, o, d, tmin, ANY ? tmax : best, &t, &u, &v)) continue;
        int prim = __float_as_int(a.w);
        if (ANY) {
          h->t = t; h->u = u; h->v = v; h->prim = prim;
          if (node_visits) { *node_visits += nv; *tri_tests += nt; }
          return true;
        }
        if (!found || t < best) {
          best = t; h->t = t; h->u = u; h->v = v; h->prim = prim; h->tie = 0; found = true;
        } else if (t == best) {
          h->tie = 1;
          if (prim < h->prim) { h->u = u; h->v = v; h->prim = prim; }
        }
      }
    }
    if (sp == 0) break;
    node = stk[stride * --sp];
  }
  if (node_visits) { *node_visits += nv; *tri_tests += nt; }
  return found;
}

// Hit reconstruction, core/kernel.cpp:127-172.
struct Isect {
  V3 p, ng, ns, u, v;
  int prim, mat, light;
};
__device__ __forceinline__ void load_prim(const DScene &sc, int prim, V3 *p0, V3 *p1, V3 *p2, V3 *n0, V3 *n1,
                                          V3 *n2, int *mat, int *light) {
  const float4 *s = sc.shade + 5 * (size_t)prim;
  float4 a = s[0], b = s[1], c = s[2], d = s[3], e = s[4];
  *p0 = mk(a.x, a.y, a.z); *p1 = mk(a.w, b.x, b.y); *p2 = mk(b.z, b.w, c.x);
  *n0 = mk(c.y, c.z, c.w); *n1 = mk(d.x, d.y, d.z); *n2 = mk(d.w, e.x, e.y);
  *mat = __float_as_int(e.z);
  *light = __float_as_int(e.w);
}
__device__ inline void make_isect(const DScene &sc, int prim, float bu, float bv, Isect *is) {
  V3 p0, p1, p2, n0, n1, n2;
  load_prim(sc, prim, &p0, &p1, &p2, &n0, &n1, &n2, &is->mat, &is->light);
  float w = 1 - bu - bv;
  is->p = add(add(mul(p0, w), mul(p1, bu)), mul(p2, bv));
  is->ns = add(add(mul(n0, w), mul(n1, bu)), mul(n2, bv));
  is->ng = normalize(cross(sub(p1, p0), sub(p2, p0)));
  V3 u = normalize(sub(p1, p0));
  V3 v = normalize(cross(is->ns, u));
  u = normalize(cross(v, is->ns));
  is->u = u;
  is->v = v;
  is->prim = prim;
}

// Sobol (sobol.h:28-43): XOR of generator columns for the set bits of the
// index, float(result) * 2^-32.  Indices on this path are < 2^32.
__device__ __forceinline__ float sobol(const uint32_t *tab, uint32_t index, int dim) {
  const uint32_t *m = tab + (size_t)dim * 32;
  uint32_t r = 0;
  while (index) {
    int b = __builtin_ctz(index);
    r ^= m[b];
    index &= index - 1;
  }
  return (float)r * (1.f / 4294967296.0f);
}

}  // namespace dev
}  // namespace orion_amd
