// Device scene layout in HBM and the BVH traversal (replaces rtcIntersect,
// core/kernel.cpp:111-174).
//
// Closest-hit contract (shared with the CPU side): Moller-Trumbore in float
// on (p0, e1 = p1 - p0, e2 = p2 - p0), its cross and dot products with
// explicit fused multiply-adds and nothing else contracted; a triangle is
// accepted when tmin < t <= tmax; among accepted triangles the hit is the
// lexicographic minimum of (t, primitive id), so the answer does not depend
// on traversal order.  `tie` flags rays where a second triangle was accepted
// at exactly the winning t (shared edges / vertices).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/orion_amd.h"
#include "bvh.h"
#include "device_math.h"

namespace orion_amd {

struct DScene {
  const float4 *nodes;   // 4 x float4 per node (bvh.h)
  const float4 *tris;    // 3 x float4 per leaf slot
  const float4 *shade;   // 5 x float4 per primitive: p0 p1 p2 n0 n1 n2 (mat, light bits)
  const float *mats;     // ORION_MAT_FLOATS floats per material
  const float *lights;   // 32 floats per light
  const float *light_cdf;   // n_lights + 1
  const float *light_func;  // n_lights
  const float *light_area;  // n_lights (area lights; 0 otherwise)
  const uint32_t *sobol;    // 1024 dims x 4 bytes x 256: Sobol byte tables
  const float *env_tex;     // envmap texels x L, RGB
  const float *env_dist;    // envmap Distribution2D tables
  const int *env_meta;      // per envmap: light, width, height, texel offset, table offset
  int n_env;                // envmap lights, in light order
  float world_c[3], world_r;  // Scene::worldbound().boundingSphere
  float light_func_int;
  int n_lights;
  int n_prims;
  int stack_bound;    // traversal stack entries that can be live at once
  float cam[16];
};

namespace dev {

// Robust slab test (Ize 2013 / pbrt-v3) on a box given by its near and far
// planes along the ray (per axis: the lower plane when the direction's
// component is positive, the upper one otherwise).  The exit distance is
// scaled by 1 + 2 gamma(3) so rounding never culls a box the exact test would
// enter; the scale is applied once to min(t_far) -- identical to scaling each
// slab first, as multiplication by a positive constant is monotonic.  Taking
// the planes by the direction's sign gives the same entry and exit distances
// as min / max over both planes of each axis: (plane - o) * inv is monotonic
// in the plane (rounded subtraction and multiplication preserve order), and
// a box's lower plane never exceeds its upper one.
__device__ __forceinline__ bool box_hit(float nx, float fx, float ny, float fy, float nz, float fz, V3 o, V3 inv,
                                        float tmin, float tmax, float *tnear) {
  const float anx = (nx - o.x) * inv.x, afx = (fx - o.x) * inv.x;
  const float any = (ny - o.y) * inv.y, afy = (fy - o.y) * inv.y;
  const float anz = (nz - o.z) * inv.z, afz = (fz - o.z) * inv.z;
  const float t0 = fmaxf(fmaxf(anx, any), fmaxf(anz, tmin));
  const float t1 = fminf(fminf(afx, afy), afz) * 1.00000035762786865234375f;
  *tnear = t0;
  return t0 <= fminf(t1, tmax);
}

// Cross and dot products of the triangle test with explicit fused
// multiply-adds (each fmaf rounds once, identically on the CPU): 2 and 3
// instructions where the unfused forms take 3 and 5.
__device__ __forceinline__ V3 cross_fma(V3 a, V3 b) {
  return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float dot_fma(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// Triangle test: identical operation sequence to the CPU restatement.
__device__ __forceinline__ bool tri_test(V3 p0, V3 e1, V3 e2, V3 o, V3 d, float tmin, float tmax, float *t,
                                         float *u, float *v) {
  V3 pv = cross_fma(d, e2);
  float det = dot_fma(e1, pv);
  if (det == 0.0f || det != det) return false;
  float inv = rcp_rn(det);
  V3 tv = sub(o, p0);
  float uu = dot_fma(tv, pv) * inv;
  if (!(uu >= 0.0f && uu <= 1.0f)) return false;
  V3 qv = cross_fma(tv, e1);
  float vv = dot_fma(d, qv) * inv;
  if (!(vv >= 0.0f && uu + vv <= 1.0f)) return false;
  float tt = dot_fma(e2, qv) * inv;
  if (!(tt > tmin && tt <= tmax)) return false;
  *t = tt; *u = uu; *v = vv;
  return true;
}

struct HitRec {
  float t, u, v;
  int prim;
  int tie;
};

// Traversal stack: the newest RING entries of each lane live in an LDS ring
// (`ring` = the lane's first slot, slots kRingStride ints apart, one LDS
// column per lane so a wave's accesses are conflict-free); older entries
// move to the wave's spill area in HBM (`spill`, wave-uniform: lane l's entry
// e is spill[e * spill_stride + l], `spill_cap` entries; the offset is formed
// in 64 bits: a 2^26-path wave's launch has up to 2^27 threads).  A DFS stack
// of a 4-wide tree holds at most three siblings per level, so stack_bound
// entries never overflow; an overflow (a malformed tree) is counted in
// *overflow, the entry dropped, and the host fails the call with
// ORION_ERR_UNSUPPORTED.
constexpr int kRingStride = 64;  // threads per traversal block: one wave64 (2 and 4 measured 3% and 20% slower)
// Per-lane LDS words after the ring: values the engine needs only when a
// triangle is accepted or a ray is stored (the closest hit's u and v, the
// ray's queue tag) live there instead of in registers.
constexpr int kLaneStash = 3;
enum { STASH_U = 0, STASH_V = 1, STASH_TAG = 2 };
// LDS ints per traversal block for a ring of RING entries per lane
template <int RING>
constexpr int trace_lds_ints() { return (RING + kLaneStash) * kRingStride; }

template <int RING>
struct TravStack {
  int *ring;
  int *spill;  // the wave's spill area (scalar)
  int spill_stride, spill_cap;
  int top = 0, base = 0;  // the ring holds [base, top), the spill area [0, base)
  __device__ __forceinline__ void reset() { top = base = 0; }
  __device__ __forceinline__ void stash(int j, int x) { ring[kRingStride * (RING + j)] = x; }
  __device__ __forceinline__ int unstash(int j) const { return ring[kRingStride * (RING + j)]; }
  __device__ __forceinline__ int *spill_slot(int e) const {
    return spill + ((size_t)spill_stride * e + __lane_id());
  }
  __device__ __forceinline__ void push(int x, uint32_t *overflow) {
    if (top - base == RING) {
      if (base < spill_cap) *spill_slot(base) = ring[kRingStride * (base & (RING - 1))];
      else ++*overflow;
      ++base;
    }
    ring[kRingStride * (top & (RING - 1))] = x;
    ++top;
  }
  __device__ __forceinline__ int pop() {
    if (top > base) {
      --top;
      return ring[kRingStride * (top & (RING - 1))];
    }
    if (base > 0) {
      --base;
      --top;
      return base < spill_cap ? *spill_slot(base) : kLeafEmpty;
    }
    return kLeafEmpty;
  }
};

__device__ __forceinline__ float q2f(uint32_t word, int k) { return (float)((word >> (8 * k)) & 0xffu); }

// One ray in flight in the traversal engine.  Any-hit rays: a triangle at
// t < tmax occludes; one at exactly tmax only when its primitive id is below
// `tie` (kTieAll: every one), so a visibility query on (0, t_L] towards
// primitive L with tie = L answers "is L the closest hit" in the closest-hit
// order, lexicographic in (t, primitive id).
constexpr int kTieAll = 0x7fffffff;
struct TRay {
  V3 o, d, inv;
  float tmin, tmax;
  int tag;  // caller's id (queue entry)
  int tie;
};

// |d| components below 1e-35 are clamped so (lo - o) * inv is never 0 * inf
// (a NaN would drop a slab); the padded boxes absorb the change.
__device__ __forceinline__ V3 safe_inv(V3 d) {
  auto safe = [](float x) { return fabsf(x) > 1e-35f ? x : copysignf(1e-35f, x); };
  return mk(rcp_rn(safe(d.x)), rcp_rn(safe(d.y)), rcp_rn(safe(d.z)));
}

// Inner node: decode the four child boxes (fmaf(q, step, origin), exactly as
// the builder evaluated them), slab-test them, continue with the nearest hit
// child and push the others far-to-near.  Returns the next node (a leaf code
// or kLeafEmpty when the stack is exhausted).
template <bool ANY, int RING>
__device__ __forceinline__ int node_step(const TRay &ray, float best, float4 A, float4 f1, float4 f2, uint2 f3,
                                         TravStack<RING> &st, uint32_t *overflow) {
  const int4 C = make_int4(__float_as_int(f1.x), __float_as_int(f1.y), __float_as_int(f1.z), __float_as_int(f1.w));
  const uint4 Q = make_uint4(__float_as_uint(f2.x), __float_as_uint(f2.y), __float_as_uint(f2.z), __float_as_uint(f2.w));
  const uint2 R = f3;
  const uint32_t eb = __float_as_uint(A.w);
  const float sx = __uint_as_float((eb & 0xffu) << 23);
  const float sy = __uint_as_float(((eb >> 8) & 0xffu) << 23);
  const float sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
  // the four children's near and far plane bytes per axis, chosen once per
  // node by the signs of the ray's inverse direction
  const bool negx = ray.inv.x < 0.0f, negy = ray.inv.y < 0.0f, negz = ray.inv.z < 0.0f;
  const uint32_t nxw = negx ? Q.y : Q.x, fxw = negx ? Q.x : Q.y;
  const uint32_t nyw = negy ? Q.w : Q.z, fyw = negy ? Q.z : Q.w;
  const uint32_t nzw = negz ? R.y : R.x, fzw = negz ? R.x : R.y;
  const int cc[4] = {C.x, C.y, C.z, C.w};
  float tk[4];
  int ck[4];
  int nh = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float nx = fmaf(q2f(nxw, k), sx, A.x), fx = fmaf(q2f(fxw, k), sx, A.x);
    float ny = fmaf(q2f(nyw, k), sy, A.y), fy = fmaf(q2f(fyw, k), sy, A.y);
    float nz = fmaf(q2f(nzw, k), sz, A.z), fz = fmaf(q2f(fzw, k), sz, A.z);
    float t;
    bool hit = box_hit(nx, fx, ny, fy, nz, fz, ray.o, ray.inv, ray.tmin, best, &t) & (cc[k] != kLeafEmpty);
    tk[k] = hit ? t : __builtin_inff();
    ck[k] = cc[k];
    nh += hit ? 1 : 0;
  }
  auto cs = [&](int a, int b) {
    // ascending entry distance (missed children last), for any-hit rays too:
    // measured 2% faster than visiting hit children in slot order (a minimum
    // tree for the nearest child with the others pushed unsorted: 0.5%
    // slower for any-hit rays)
    bool sw = tk[b] < tk[a];
    float ta = tk[a], tb = tk[b];
    int ca = ck[a], cb = ck[b];
    tk[a] = sw ? tb : ta; tk[b] = sw ? ta : tb;
    ck[a] = sw ? cb : ca; ck[b] = sw ? ca : cb;
  };
  cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
  if (nh > 3) st.push(ck[3], overflow);
  if (nh > 2) st.push(ck[2], overflow);
  if (nh > 1) st.push(ck[1], overflow);
  return nh > 0 ? ck[0] : st.pop();
}

// One triangle of a leaf (its three float4s): the triangle test, keeping the
// lexicographic minimum (t, prim) and the tie flag.  Any-hit returns true
// when accepted.
template <bool ANY, class ST>
__device__ __forceinline__ bool tri_step(const TRay &ray, float4 a, V3 e1, V3 e2, float *best, HitRec *h,
                                         bool *found, ST &st) {
  float t, u, v;
  if (!tri_test(mk(a.x, a.y, a.z), e1, e2, ray.o, ray.d, ray.tmin, ANY ? ray.tmax : *best, &t, &u, &v)) return false;
  int prim = __float_as_int(a.w);
  if (ANY) {
    if (!(t < ray.tmax || prim < ray.tie)) return false;
    h->t = t; h->u = u; h->v = v; h->prim = prim;
    *found = true;
    return true;
  }
  // closest hit: its t is *best, its u and v wait in the lane's LDS words
  // (h->t, h->u, h->v are filled in when the ray is stored)
  if (!*found || t < *best) {
    *best = t; h->prim = prim; h->tie = 0; *found = true;
    st.stash(STASH_U, __float_as_int(u));
    st.stash(STASH_V, __float_as_int(v));
  } else if (t == *best) {
    h->tie = 1;
    if (prim < h->prim) {
      h->prim = prim;
      st.stash(STASH_U, __float_as_int(u));
      st.stash(STASH_V, __float_as_int(v));
    }
  }
  return false;
}

// Traversal engine for one wave over queue entries [begin, end).  Every
// iteration each lane does up to one inner-node visit and one triangle of
// its open leaf (the next node is popped when a leaf opens, so a lane often
// has both; their fetches are in flight together: measured 6% faster
// traversal than one unit per iteration, at 6 waves/SIMD instead of 8), so
// lanes never wait for the slowest lane to reach a leaf (with a
// descend-then-intersect loop a wave64 kept ~16% of its lanes busy).  A node
// visit that reaches a second leaf while one is open keeps it until the
// open one is done.  A lane whose ray finishes takes the next entry once at
// least kRefillMin lanes are idle.  Waves own fixed ranges (no shared work
// counter: device-scope atomics on one address serialise across the XCDs
// and cost more than they balance).
//
// A refill costs no more memory latency than any other iteration: the queue
// values of the next refill's entries are fetched ahead (lane l holds entry
// next + l's and hands it to the idle lane of rank l), and a new ray's
// fields are first used after the iteration's node fetches are issued, so
// its loads and its root node's are in flight together.
//
// IO: int fetch(int entry) -> the entry's queue value; void load(int value,
// TRay *ray) (ray->tag = value); void store(const TRay &, bool hit, const
// HitRec &); nodes, tris: the scene, the same for every ray of the call
// (kept in scalar registers).
constexpr int kRefillMin = 8;

template <bool ANY, int RING, class IO>
__device__ inline void trace_engine(IO &io, int begin, int end, TravStack<RING> &st, uint32_t *node_visits,
                                    uint32_t *tri_tests, uint32_t *overflow) {
  const float4 *const nodes = io.nodes, *const tris = io.tris;
  int next = begin;  // wave-uniform: next unassigned entry
  bool active = false;
  TRay ray;
  HitRec h;
  h.prim = -1;
  h.tie = 0;
  bool found = false;
  float best = 0.f;
  int node = kLeafEmpty;  // next inner node / leaf to visit
  int tcur = 0, tend = 0;  // triangles of the open leaf
  uint32_t nv = 0, nt = 0;
  bool fresh = false;  // the lane took a new ray this iteration (its fields are still in flight)
  // queue value of entry next + lane: what the idle lane of rank `lane`
  // takes at the next refill
  const int lane = (int)__lane_id();
  int pre = next + lane < end ? io.fetch(next + lane) : 0;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const int nidle = __popcll(idle);
    if (next < end && (nidle >= kRefillMin || nidle == 64)) {
      // the lane's rank among the idle lanes below it, and that entry's
      // queue value from the lane that fetched it
      const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
      const int val = __shfl(pre, rank);
      if (!active && next + rank < end) {
        io.load(val, &ray);
        st.stash(STASH_TAG, val);
        fresh = true;
        active = true;
        found = false;
        h.prim = -1;
        h.tie = 0;
        node = 0;
        tcur = tend = 0;
        st.reset();
      }
      next = min(next + nidle, end);
      pre = next + lane < end ? io.fetch(next + lane) : 0;
    }
    if (__ballot(active) == 0) break;
    // a lane with an open leaf and a popped inner node does one of each per
    // iteration: both fetches are in flight together
    const bool do_tri = tcur < tend, do_node = node >= 0;
    float4 f0, f1, f2, g0;
    uint2 f3;
    V3 e1, e2;
    if (do_tri) {
      const float4 *src = tris + 3 * (size_t)tcur;
      g0 = gld(src);
      e1 = gld3((const float *)(src + 1));
      e2 = gld3((const float *)(src + 2));
    }
    if (do_node) {
      const float4 *src = nodes + 4 * (size_t)node;
      f0 = gld(src); f1 = gld(src + 1); f2 = gld(src + 2); f3 = gld((const uint2 *)(src + 3));
    }
    if (fresh) {  // the new ray's fields, once its root node's fetch is issued
      ray.inv = safe_inv(ray.d);
      best = ray.tmax;
      fresh = false;
    }
    // visit counters are kept per wave (scalar registers)
    nt += (uint32_t)__popcll(__ballot(do_tri));
    nv += (uint32_t)__popcll(__ballot(do_node));
    if (do_tri) {
      ++tcur;
      if (tri_step<ANY>(ray, g0, e1, e2, &best, &h, &found, st)) {
        tcur = tend;  // any-hit: occluded, done
        node = kLeafEmpty;
      }
    }
    if (do_node && node >= 0) node = node_step<ANY, RING>(ray, best, f0, f1, f2, f3, st, overflow);
    if (tcur >= tend && node < 0 && node != kLeafEmpty) {  // open the leaf reached
      const int code = ~node;  // first triangle << 3 | count - 1
      tcur = code >> 3;
      tend = tcur + (code & 7) + 1;
      node = st.pop();
    }
    if (active && tcur >= tend && node == kLeafEmpty) {
      ray.tag = st.unstash(STASH_TAG);
      if (!ANY && found) {
        h.t = best;
        h.u = __int_as_float(st.unstash(STASH_U));
        h.v = __int_as_float(st.unstash(STASH_V));
      }
      io.store(ray, found, h);
      active = false;
    }
  }
  *node_visits += nv;  // the wave's totals, the same in every lane
  *tri_tests += nt;
}

// Hit reconstruction, core/kernel.cpp:127-172.
struct Isect {
  V3 p, ng, ns, u, v;
  int prim, mat, light;
};
__device__ __forceinline__ void load_prim(const DScene &sc, int prim, V3 *p0, V3 *p1, V3 *p2, V3 *n0, V3 *n1,
                                          V3 *n2, int *mat, int *light) {
  const float4 *s = sc.shade + 5 * (size_t)prim;
  float4 a = gld(s), b = gld(s + 1), c = gld(s + 2), d = gld(s + 3), e = gld(s + 4);
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
// index, float(result) * 2^-32.  Indices on this path are < 2^32.  `tab`
// holds per-dimension byte tables (4 x 256 words: the XOR of the columns of
// each byte value), so the XOR over the set bits is four independent loads.
__device__ __forceinline__ float sobol(const uint32_t *tab, uint32_t index, int dim) {
  const uint32_t *m = tab + (size_t)dim * 1024;
  uint32_t r = gld(m + (index & 0xffu)) ^ gld(m + 256 + ((index >> 8) & 0xffu)) ^
               gld(m + 512 + ((index >> 16) & 0xffu)) ^ gld(m + 768 + (index >> 24));
  return (float)r * (1.f / 4294967296.0f);
}

}  // namespace dev
}  // namespace orion_amd
