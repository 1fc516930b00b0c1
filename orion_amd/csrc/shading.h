// Device shading shared by the path tracer and SPPM: camera rays, light
// selection and sampling, and the deferred direct-lighting estimate
// (uniformSampleOneLight + estimateDirect, core/integrator.cpp:8-90) split
// into a setup half (before the shadow / MIS rays are traced) and a resolve
// half (after).
#pragma once
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "wavefront.h"

namespace orion_amd {
namespace dev {

// pinhole.cpp:8-33; next2() = (x: dim d+1, y: dim d) as g++ builds
// Point2f(next(), next()) (sampler.h:31).
__device__ __forceinline__ void camera_ray(const DScene &sc, int i, int j, uint32_t index, V3 *o, V3 *d) {
  float uy = sobol(sc.sobol, index, 5);
  float ux = sobol(sc.sobol, index, 6);
  float rx = (float)i + ux, ry = (float)j + uy;
  const float *c = sc.cam;
  float width = c[14], height = c[15];
  float xx = (2 * rx / width - 1) * c[13] * c[12];
  float yy = (1 - 2 * ry / height) * c[12];
  V3 dir = normalize(mk(xx, yy, -1.0f));
  *o = mk(c[0], c[1], c[2]);
  *d = mat3(c + 3, dir);
}

// Distribution1D::sampleDiscrete (sampling.h:110-120) over light power.
__device__ __forceinline__ int sample_light(const DScene &sc, float u, float *pdf) {
  int size = sc.n_lights + 1;
  int first = 0, len = size;
  while (len > 0) {
    int half = len >> 1, mid = first + half;
    if (gld(sc.light_cdf + mid) <= u) { first = mid + 1; len -= half + 1; }
    else len = half;
  }
  int off = first - 1;
  off = off < 0 ? 0 : (off > size - 2 ? size - 2 : off);
  *pdf = (sc.light_func_int > 0) ? gld(sc.light_func + off) / (sc.light_func_int * sc.n_lights) : 0;
  return off;
}

__device__ __forceinline__ float spot_falloff(const float *L, V3 w) {  // spot.cpp:10-19
  V3 wl = normalize(mat3(L + 11, w));
  float cosTheta = wl.z;
  if (cosTheta < gld(L + 9)) return 0;
  if (cosTheta > gld(L + 10)) return 1;
  float delta = (cosTheta - gld(L + 9)) / (gld(L + 10) - gld(L + 9));
  return (delta * delta) * (delta * delta);
}

// Triangle::sample (triangle.cpp:85-106): point, face-forwarded ng
__device__ __forceinline__ void tri_sample(const DScene &sc, int prim, float u0, float u1, V3 *p, V3 *ng) {
  V3 p0, p1, p2, n0, n1, n2;
  int m_, l_;
  load_prim(sc, prim, &p0, &p1, &p2, &n0, &n1, &n2, &m_, &l_);
  float su0 = sqrtf(u0);
  float b0 = 1 - su0, b1 = u1 * su0;
  *p = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, 1 - b0 - b1));
  V3 g = normalize(cross(sub(p1, p0), sub(p2, p0)));
  V3 ns = add(add(mul(n0, b0), mul(n1, b1)), mul(n2, 1.0f - b0 - b1));
  *ng = dot(g, ns) < 0.f ? neg(g) : g;
}

// ---------------------------------------------------------------- infinite lights
constexpr float kInv2Pi = 0.15915494309189533577f;
enum { LT_AREA = 0, LT_POINT = 1, LT_SPOT = 2, LT_DISTANT = 3, LT_ENVMAP = 4 };

// Envmap e's image and Distribution2D tables (light/envmap.cpp:9-42,
// sampling.h:127-150).  Table layout: marginal [nv func][nv+1 cdf][funcInt],
// then per row v: [nu func][nu+1 cdf][funcInt].
struct EnvView {
  int nu, nv;
  const float *tex;
  const float *d;
  __device__ __forceinline__ const float *row(int v) const { return d + 2 * nv + 2 + (size_t)v * (2 * nu + 2); }
};
__device__ __forceinline__ EnvView env_view(const DScene &sc, int e) {
  const int *m = sc.env_meta + 5 * e;
  EnvView v;
  v.nu = gld(m + 1);
  v.nv = gld(m + 2);
  v.tex = sc.env_tex + gld(m + 3);
  v.d = sc.env_dist + gld(m + 4);
  return v;
}
// Texture::sample(float, float) then ImageTexture's wrap filter
// (texture.cpp:4-8, 13-24).  u, v lie in [0, 1] here; a negative multiple of
// the width wraps to 0 (the reference would index one past the row).
__device__ __forceinline__ Sp env_lookup(const EnvView &v, float u, float w) {
  int x = (int)(u * v.nu), y = (int)(w * v.nv);
  x %= v.nu;
  if (x < 0) x += v.nu;
  y %= v.nv;
  if (y < 0) y += v.nv;
  const float *t = v.tex + 3 * ((size_t)y * v.nu + x);
  return sp3(gld(t), gld(t + 1), gld(t + 2));
}
// Distribution1D::sampleContinuous (sampling.h:91-109) over [func][cdf][funcInt]
__device__ inline float dist_sample(const float *d, int n, float u, float *pdf, int *off) {
  const float *cdf = d + n;
  int size = n + 1, first = 0, len = size;
  while (len > 0) {
    int half = len >> 1, mid = first + half;
    if (gld(cdf + mid) <= u) { first = mid + 1; len -= half + 1; }
    else len = half;
  }
  int o = first - 1;
  o = o < 0 ? 0 : (o > size - 2 ? size - 2 : o);
  *off = o;
  float c0 = gld(cdf + o), c1 = gld(cdf + o + 1);
  float du = u - c0;
  if ((c1 - c0) > 0) du /= (c1 - c0);
  float fint = gld(d + 2 * n + 1);
  *pdf = (fint > 0) ? gld(d + o) / fint : 0;
  return (o + du) / n;
}
// Distribution2D::SampleContinuous / Pdf (sampling.h:131-145)
__device__ inline void env_sample_uv(const EnvView &v, float u0, float u1, float *su, float *sv, float *pdf) {
  float p1, p0;
  int row, c_;
  *sv = dist_sample(v.d, v.nv, u1, &p1, &row);
  *su = dist_sample(v.row(row), v.nu, u0, &p0, &c_);
  *pdf = p0 * p1;
}
__device__ inline float env_uv_pdf(const EnvView &v, float pu, float pv) {
  int iu = (int)(pu * v.nu), iv = (int)(pv * v.nv);
  iu = iu < 0 ? 0 : (iu > v.nu - 1 ? v.nu - 1 : iu);
  iv = iv < 0 ? 0 : (iv > v.nv - 1 ? v.nv - 1 : iv);
  return gld(v.row(iv) + iu) / gld(v.d + 2 * v.nv + 1);
}
// The envmap routines are kept out of line: their FP64 acos / atan2 / sin /
// cos would otherwise be inlined at every call site of the shading kernels
// and push the kernels' register allocation into scratch.
// Envmap::Le (envmap.cpp:44-49); L = the light record (world-to-light at [17])
__device__ __noinline__ Sp env_le(const DScene &sc, const float *L, V3 dir) {
  EnvView v = env_view(sc, (int)gld(L + 1));
  V3 w = normalize(mat3(L + 17, dir));
  float u = spherical_phi(w) * kInv2Pi, t = spherical_theta(w) * kInvPi;
  return env_lookup(v, u, t);
}
// Envmap::pdf_Li (envmap.cpp:74-82): phi and theta both scaled by 1 / (2 pi)
// as in the reference, and its zero-sine test compares a function address
// (never true), so the division is always taken.
__device__ __noinline__ float env_pdf_li(const DScene &sc, const float *L, V3 wi) {
  EnvView v = env_view(sc, (int)gld(L + 1));
  V3 swi = mat3(L + 17, wi);
  float theta = spherical_theta(swi), phi = spherical_phi(swi);
  float sinTh = dsin(theta);
  return env_uv_pdf(v, phi * kInv2Pi, theta * kInv2Pi) / (2 * kPi * kPi * sinTh);
}
// Sum of Le over the scene's envmaps in light order (path.cpp:65-68,
// sppm.cpp:95-99), added to L one envmap at a time.
__device__ inline void add_env_le(const DScene &sc, V3 dir, Sp beta, Sp *L) {
  for (int e = 0; e < sc.n_env; ++e) {
    const float *LR = sc.lights + 32 * (size_t)gld(sc.env_meta + 5 * e);
    acc(*L, beta * env_le(sc, LR, dir));
  }
}

// Envmap::sample_Li's direction, pdf and radiance (envmap.cpp:55-73); *pdf = 0
// where the map pdf is 0
__device__ __noinline__ Sp env_sample_li(const DScene &sc, const float *L, float u0, float u1, V3 *wi, float *pdf) {
  EnvView v = env_view(sc, (int)gld(L + 1));
  float su, sv, mapPdf;
  env_sample_uv(v, u0, u1, &su, &sv, &mapPdf);
  if (mapPdf == 0) { *pdf = 0; return sp(0); }
  float theta = sv * kPi, phi = su * 2 * kPi;
  float cosTheta = dcos(theta), sinTheta = dsin(theta);
  float sinPhi = dsin(phi), cosPhi = dcos(phi);
  *wi = mat3(L + 8, mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta));
  *pdf = mapPdf / (2 * kPi * kPi * sinTheta);
  if (sinTheta == 0) *pdf = 0;
  return env_lookup(v, su, sv);
}

// Light::sample_Li; *so / *sd = the ShadowTester ray (intersection.h:40-44)
__device__ inline Sp light_sample_li(const DScene &sc, int l, V3 refp, V3 refng, float u0, float u1, V3 *wi,
                                     float *pdf, V3 *so, V3 *sd) {
  const float *L = sc.lights + 32 * (size_t)l;
  int type = (int)gld(L + 0);
  Sp Le = sp3(gld(L + 2), gld(L + 3), gld(L + 4));
  V3 target;
  if (type == 0) {  // diffuse.cpp:13-25 + triangle.cpp:108-121
    V3 p, ng;
    tri_sample(sc, (int)gld(L + 1), u0, u1, &p, &ng);
    float pd = 1.0f / gld(sc.light_area + l);
    V3 w = sub(p, refp);
    if (lensq(w) == 0) pd = 0;
    else {
      w = normalize(w);
      pd *= lensq(sub(refp, p)) / absdot(ng, neg(w));
      if (is_inf(pd)) pd = 0;
    }
    if (pd == 0 || lensq(sub(p, refp)) == 0) { *pdf = 0; return sp(0); }
    *pdf = pd;
    *wi = normalize(sub(p, refp));
    target = p;
    Le = dot(ng, neg(*wi)) > 0 ? Le : sp(0);
  } else if (type == LT_DISTANT) {  // distant.cpp:16-27
    *wi = mk(gld(L + 5), gld(L + 6), gld(L + 7));
    *pdf = 1;
    target = add(refp, mul(*wi, 2 * sc.world_r));
  } else if (type == LT_ENVMAP) {
    Le = env_sample_li(sc, L, u0, u1, wi, pdf);
    if (*pdf == 0) return sp(0);
    target = add(refp, mul(*wi, 2 * sc.world_r));
  } else {  // point.cpp:10-19, spot.cpp:20-29
    V3 pos = mk(gld(L + 5), gld(L + 6), gld(L + 7));
    *wi = normalize(sub(pos, refp));
    *pdf = 1.0f;
    target = pos;
    if (type == 1) Le = Le / lensq(sub(pos, refp));
    else Le = Le * spot_falloff(L, neg(*wi)) / lensq(sub(pos, refp));
  }
  *so = offset_ray_origin(refp, refng, sub(target, refp));
  *sd = sub(target, *so);
  return Le;
}

// Light::sample_Le (diffuse.cpp:32-45, point.cpp:20-27, spot.cpp:30-38)
__device__ inline Sp light_sample_le(const DScene &sc, int l, float a0, float a1, float b0, float b1, V3 *ro,
                                     V3 *rd, V3 *nLight, float *pdfPos, float *pdfDir) {
  const float *L = sc.lights + 32 * (size_t)l;
  int type = (int)gld(L + 0);
  Sp Le = sp3(gld(L + 2), gld(L + 3), gld(L + 4));
  if (type == 0) {
    V3 p, ng;
    tri_sample(sc, (int)gld(L + 1), a0, a1, &p, &ng);
    *pdfPos = 1.0f / gld(sc.light_area + l);
    *nLight = ng;
    V3 w = cosine_hemisphere(b0, b1);
    *pdfDir = w.z * kInvPi;
    V3 x, y;
    coordinate_system(ng, &x, &y);
    w = add(add(mul(x, w.x), mul(y, w.y)), mul(ng, w.z));
    *ro = offset_ray_origin(p, ng, w);
    *rd = w;
    return dot(ng, w) > 0 ? Le : sp(0);
  }
  if (type == LT_DISTANT || type == LT_ENVMAP) {
    V3 dir;
    float cx, cy, mapPdf = 1, sinTheta = 1, su = 0, sv = 0;
    EnvView v;
    if (type == LT_DISTANT) {  // distant.cpp:29-43
      dir = mk(gld(L + 5), gld(L + 6), gld(L + 7));
      concentric_disk(a0, a1, &cx, &cy);
    } else {  // envmap.cpp:83-107
      v = env_view(sc, (int)gld(L + 1));
      env_sample_uv(v, a0, a1, &su, &sv, &mapPdf);
      if (mapPdf == 0) { *pdfPos = 0; *pdfDir = 0; return sp(0); }
      float theta = sv * kPi, phi = su * 2.f * kPi;
      float cosTheta = dcos(theta);
      sinTheta = dsin(theta);
      float sinPhi = dsin(phi), cosPhi = dcos(phi);
      dir = mat3(L + 8, mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta));  // -d
      concentric_disk(b0, b1, &cx, &cy);
    }
    V3 v1, v2;
    coordinate_system(dir, &v1, &v2);
    const float r = sc.world_r;
    V3 pDisk = add(mk(sc.world_c[0], sc.world_c[1], sc.world_c[2]), mul(add(mul(v1, cx), mul(v2, cy)), r));
    *ro = add(pDisk, mul(dir, r));
    *rd = neg(dir);
    *nLight = *rd;
    *pdfPos = 1 / (kPi * r * r);
    if (type == LT_DISTANT) {
      *pdfDir = 1;
      return Le;
    }
    *pdfDir = sinTheta == 0 ? 0 : mapPdf / (2 * kPi * kPi * sinTheta);
    return env_lookup(v, su, sv);
  }
  V3 pos = mk(gld(L + 5), gld(L + 6), gld(L + 7));
  *ro = pos;
  *pdfPos = 1;
  if (type == 1) {
    *rd = uniform_sphere(a0, a1);
    *nLight = *rd;
    *pdfDir = kInv4Pi;
    return Le;
  }
  V3 w = uniform_cone(a0, a1, gld(L + 9));
  *rd = mat3(L + 20, w);
  *nLight = *rd;
  *pdfDir = 1 / (2 * kPi * (1 - gld(L + 9)));
  return Le * spot_falloff(L, *rd);
}

// Second half of estimateDirect: the pending light-sampled term (when its
// shadow ray was unoccluded) plus the BSDF-sampled term (when its
// visibility ray was), as the radiance term beta * (Ld / lightPdf).
__device__ inline Sp direct_term(const Wave &w, int p, int fl) {
  Sp Ld = sp(0);
  const float4 A = w.pa[p];
  if ((fl & F_SHADOW) && w.vis[p]) acc(Ld, sp3(A.x, A.y, A.z));
  if ((fl & F_MIS) && w.vis[w.N + p]) {
    const float4 M = w.pm[p];
    acc(Ld, sp3(M.x, M.y, M.z));
  }
  const float4 nb = w.pn[p];
  return sp3(nb.x, nb.y, nb.z) * (Ld / A.w);
}
__device__ inline void resolve_direct(const Wave &w, int p, int fl, Sp *L) { acc(*L, direct_term(w, p, fl)); }

// uniformSampleOneLight's light choice (core/integrator.cpp:8-16): by the
// power CDF, or uniform when `uniform`; one Sobol draw.  Returns the light,
// or -1 when there is nothing to sample: no lights (no draw) or a zero
// selection pdf (the reference returns after the draw).
__device__ inline int nee_pick(const DScene &sc, uint32_t index, int *dim, bool uniform, float *sel) {
  if (sc.n_lights == 0) return -1;
  if (!uniform) {
    const int l = sample_light(sc, sobol(sc.sobol, index, (*dim)++), sel);
    return *sel == 0 ? -1 : l;
  }
  const int nl = sc.n_lights;
  *sel = 1.0f / nl;
  return min((int)(sobol(sc.sobol, index, (*dim)++) * nl), nl - 1);
}

// The traversal's triangle test of primitive `prim` (tri_test on the leaf
// record p0, p1 - p0, p2 - p0, the differences formed in float as the BVH
// builder forms them) for the ray (o, d) on (0, FLT_MAX]: the t, u, v a
// closest-hit traversal reports when `prim` is the hit.
__device__ inline bool prim_hit(const DScene &sc, int prim, V3 o, V3 d, float *t, float *u, float *v) {
  const float4 *s = sc.shade + 5 * (size_t)prim;
  const float4 a = gld(s), b = gld(s + 1), c = gld(s + 2);
  const V3 p0 = mk(a.x, a.y, a.z);
  const V3 e1 = mk(a.w - a.x, b.x - a.y, b.y - a.z);
  const V3 e2 = mk(b.z - a.x, b.w - a.y, c.x - a.z);
  return tri_test(p0, e1, e2, o, d, 0.0f, kFltMax, t, u, v);
}

// estimateDirect's sampling half (core/integrator.cpp:29-90) for light l,
// chosen with probability sel: light and BSDF samples (four draws), the
// light-sampled term A with its shadow ray (slot 2), and the BSDF-sampled
// term M with its visibility ray (slot 1).  orion counts M when the
// BSDF-sampled ray's closest hit is the light (integrator.cpp:66-88).  For
// an area light that is the light's triangle: it is tested here with the
// traversal's own test and, when hit at t_L, M is evaluated at that hit and
// the ray becomes a visibility query on (0, t_L] with tie = the triangle
// (a triangle at exactly t_L occludes only with a lower primitive id, the
// closest-hit order); a direction that misses the triangle has no M and no
// ray.  For an envmap M counts when the ray escapes: a visibility query on
// (0, FLT_MAX].  Returns the F_PEND | F_SHADOW | F_MIS bits; *ref_rays gets
// the closest-hit casts orion makes for the BSDF sample (0 or 1).
__device__ inline int nee_sample(const DScene &sc, const Wave &w, int p, int l, float sel, const Bsdf &bs,
                                 const Isect &is, V3 d, Sp beta, uint32_t index, int *dim, int *ref_rays) {
  const int N = w.N;
  float lu1 = sobol(sc.sobol, index, (*dim)++), lu0 = sobol(sc.sobol, index, (*dim)++);
  float bu1 = sobol(sc.sobol, index, (*dim)++), bu0 = sobol(sc.sobol, index, (*dim)++);
  const int nonspec = ALL & ~SPEC;
  V3 wo = normalize(neg(d));
  V3 wi = mk(0, 0, 0), so, sdir;
  float lpdf = 0;
  Sp Li = light_sample_li(sc, l, is.p, is.ng, lu0, lu1, &wi, &lpdf, &so, &sdir);
  const float *LR = sc.lights + 32 * (size_t)l;
  const int ltype = (int)gld(LR);
  bool delta = ltype != LT_AREA && ltype != LT_ENVMAP;  // isDeltaLight (light.h:39-42)
  bool hs = false, hm = false;
  Sp A = sp(0), M = sp(0);
  if (lpdf > 0 && !black(Li)) {
    Sp f = bsdf_f(bs, wi, wo, nonspec) * absdot(wi, is.ns);
    float bpdf = bsdf_pdf(bs, wi, wo, nonspec);
    if (!black(f)) {
      if (delta) A = f * Li / lpdf;
      else A = f * Li * power_heuristic(lpdf, bpdf) / lpdf;
      put_ray(w, 2 * (size_t)N + p, so, sdir, 1 - kShadowEps);
      hs = true;
    }
  }
  if (!delta) {
    float bp2 = 0;
    int st = 0;
    V3 wi2 = wi;
    Sp f2 = bsdf_sample(bs, &wi2, wo, bu0, bu1, &bp2, nonspec, &st);
    mulby(f2, sp(absdot(wi2, is.ns)));
    // an envmap's pdf_Li needs no ray: a zero ends the estimate before it
    bool traced = !black(f2) && bp2 > 0;
    if (traced && ltype == LT_ENVMAP && !(st & SPEC) && env_pdf_li(sc, LR, wi2) == 0) traced = false;
    if (traced) {
      ++*ref_rays;
      V3 o2 = offset_ray_origin(is.p, is.ng, wi2);
      float tvis = kFltMax;
      int tie = kTieAll;
      if (ltype == LT_ENVMAP) {
        float weight = power_heuristic(bp2, env_pdf_li(sc, LR, wi2));
        Sp Lm = env_le(sc, LR, wi2);
        if (!black(Lm)) M = f2 * Lm * weight / bp2;
      } else {
        // Triangle::pdf (triangle.cpp:123-135) at the light's hit
        const int pl = (int)gld(LR + 1);
        float tl, ul, vl;
        if (prim_hit(sc, pl, o2, wi2, &tl, &ul, &vl)) {
          Isect in;
          make_isect(sc, pl, ul, vl, &in);
          float hpdf = lensq(sub(is.p, in.p)) / (absdot(in.ng, neg(wi2)) * gld(sc.light_area + l));
          if (is_inf(hpdf)) hpdf = 0;
          if (hpdf != 0) {
            float weight = power_heuristic(bp2, hpdf);
            Sp Lm = dot(in.ng, neg(wi2)) > 0 ? sp3(gld(LR + 2), gld(LR + 3), gld(LR + 4)) : sp(0);
            if (!black(Lm)) M = f2 * Lm * weight / bp2;
          }
          tvis = tl;
          tie = pl;
        }
      }
      if (!black(M)) {
        put_ray(w, (size_t)N + p, o2, wi2, tvis, tie);
        hm = true;
      }
    }
  }
  if (!(hs || hm)) return 0;
  w.pa[p] = make_float4(A.r, A.g, A.b, sel);
  w.pm[p] = make_float4(M.r, M.g, M.b, 0.f);
  w.pn[p] = make_float4(beta.r, beta.g, beta.b, 0.f);
  return F_PEND | (hs ? F_SHADOW : 0) | (hm ? F_MIS : 0);
}

// Both halves in one place (the SPPM camera pass).  Draws exactly the
// reference's dimensions.
__device__ inline int setup_direct(const DScene &sc, const Wave &w, int p, const Bsdf &bs, const Isect &is, V3 d,
                                   Sp beta, uint32_t index, int *dim, bool uniform, int *ref_rays) {
  float sel;
  const int l = nee_pick(sc, index, dim, uniform, &sel);
  if (l < 0) return 0;
  return nee_sample(sc, w, p, l, sel, bs, is, d, beta, index, dim, ref_rays);
}

__device__ __forceinline__ Sp area_emission(const DScene &sc, const Isect &is, V3 w) {  // diffuse.cpp:53-56
  const float *LR = sc.lights + 32 * (size_t)is.light;
  return dot(is.ng, w) > 0 ? sp3(gld(LR + 2), gld(LR + 3), gld(LR + 4)) : sp(0);
}

}  // namespace dev
}  // namespace orion_amd
