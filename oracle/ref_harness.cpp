// C-ABI harness over the reference's own compiled sources (oracle/_ref).
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the translation
// units under /root/reference/src that compile and link without Embree
// (sampler/, bsdf/, core/bsdf, core/spectrum, core/transform, material/
// matte|metal|glass|mirror|plastic|uber, light/point|spot|distant|envmap,
// texture/imagetexture + the vendored stb_image_aug.c, camera/pinhole, texture/, util/strutil);
// core/triangle.cpp and light/diffuse.cpp are excluded because they need
// Scene::intersect -> EmbreeKernel, whose library ships only as a Windows
// binary.  The harness reads private members through `#define private public`
// (layout is unchanged) to report the values the reference computes.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <list>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include <xmmintrin.h>
#include <pmmintrin.h>
#include <unistd.h>
#include <stb_image_aug.h>
// Standard headers are included first so the access override below touches
// only the reference's own classes.
#define private public
#define protected public
#include <orion.h>
#include <sampler/sobol.h>
#include <sampler/sampling.h>
#include <core/transform.h>
#include <core/intersection.h>
#include <core/bsdf.h>
#include <core/spectrum.h>
#include <camera/pinhole.h>
#include <material/matte.h>
#include <material/metal.h>
#include <material/glass.h>
#include <material/mirror.h>
#include <material/plastic.h>
#include <material/uber.h>
#include <texture/constant.h>
#include <texture/floattexture.h>
#include <light/point.h>
#include <light/spot.h>
#include <light/distant.h>
#include <light/envmap.h>
#include <texture/imagetexture.h>
#include <texture/rendertarget.h>
#include <bsdf/microfacet.h>
#include <util/param.h>
#include <util/strutil.h>
#undef private
#undef protected

namespace orion {
Spectrum frConductor(Float cosThetaI, const Spectrum &etai, const Spectrum &etat, const Spectrum &k);
Float frDielectric(Float cosThetaI, Float etaI, Float etaT);
}

using namespace orion;

namespace {
void put3(float *o, const Spectrum &s) { o[0] = s.r; o[1] = s.g; o[2] = s.b; }
void put3(float *o, const Vector3f &s) { o[0] = s.x; o[1] = s.y; o[2] = s.z; }
void put3(float *o, const Point3f &s) { o[0] = s.x; o[1] = s.y; o[2] = s.z; }
void put3(float *o, const Normal3f &s) { o[0] = s.x; o[1] = s.y; o[2] = s.z; }
Vector3f v3(const float *a) { return Vector3f(a[0], a[1], a[2]); }

Intersection make_isect(const float *frame) {
  // frame: p(3) ng(3) ns(3) u(3) v(3)
  Intersection is(Point3f(frame[0], frame[1], frame[2]));
  is.ng = Normal3f(frame[3], frame[4], frame[5]);
  is.ns = Normal3f(frame[6], frame[7], frame[8]);
  is.u = Vector3f(frame[9], frame[10], frame[11]);
  is.v = Vector3f(frame[12], frame[13], frame[14]);
  is.primitive = nullptr;
  return is;
}

// Parameters are a material record without its type word:
// kind 0 matte(Kd, sigma), 1 metal(eta, k, rough) or metal preset by name,
// 2 glass(Kr, Kt, rough, -, eta), 3 mirror(Kr), 4 plastic(Kd, Ks, rough),
// 5 uber(Kd, Ks, Kr, Kt, opacity, rough, -, eta)
std::shared_ptr<ConstantTexture> ctex(const float *p) {
  return std::make_shared<ConstantTexture>(Spectrum(p[0], p[1], p[2]));
}
std::shared_ptr<Material> make_material(int kind, const float *p, const char *metal_name) {
  if (kind == 0) return createMatteMaterial(ctex(p), std::make_shared<FloatTexture>(p[3]));
  if (kind == 3) return createMirrorMaterial(ctex(p));
  if (kind == 4) return createPlasticMaterial(ctex(p), ctex(p + 3), std::make_shared<FloatTexture>(p[6]));
  if (kind == 5)
    return createUberMaterial(ctex(p), ctex(p + 3), ctex(p + 6), ctex(p + 9), ctex(p + 12),
                              std::make_shared<FloatTexture>(p[15]), std::make_shared<FloatTexture>(p[17]));
  if (kind == 1) {
    auto r = std::make_shared<FloatTexture>(p[6]);
    if (metal_name && metal_name[0]) {
      ParamSet ps;
      ps.setParam("metalName", metal_name);
      return createMetalMaterial(nullptr, nullptr, r, ps);
    }
    auto eta = std::make_shared<ConstantTexture>(Spectrum(p[0], p[1], p[2]));
    auto k = std::make_shared<ConstantTexture>(Spectrum(p[3], p[4], p[5]));
    ParamSet ps;
    return createMetalMaterial(eta, k, r, ps);
  }
  auto kr = std::make_shared<ConstantTexture>(Spectrum(p[0], p[1], p[2]));
  auto kt = std::make_shared<ConstantTexture>(Spectrum(p[3], p[4], p[5]));
  auto rough = std::make_shared<FloatTexture>(p[6]);
  auto eta = std::make_shared<FloatTexture>(p[8]);
  return createGlassMaterial(kr, kt, rough, eta);
}
}  // namespace

extern "C" {

float ref_sobol(unsigned long long index, unsigned dim) { return sobol::sample(index, dim); }

// n successive next2() draws of one sampler clone: out[2*i] = x, out[2*i+1] = y
void ref_sampler_next2(long long index, int n, float *out) {
  SobolSampler base(0);
  auto s = base.clone(index);
  for (int i = 0; i < n; ++i) {
    Point2f p = s->next2();
    out[2 * i] = p.x; out[2 * i + 1] = p.y;
  }
}

void ref_warps(const float *u, float cosThetaMax, float *out) {
  Point2f r(u[0], u[1]);
  Point2f t = uniformSampleTriangle(r);
  out[0] = t.x; out[1] = t.y;
  Point2f d = uniformSampleDisk(r);
  out[2] = d.x; out[3] = d.y;
  put3(out + 4, cosineSampleHemisphere(r));
  put3(out + 7, uniformSampleSphere(r));
  put3(out + 10, uniformSampleCone(r, cosThetaMax));
}

float ref_fr_dielectric(float c, float ei, float et) { return frDielectric(c, ei, et); }
void ref_fr_conductor(float c, const float *eta, const float *k, float *out) {
  put3(out, frConductor(c, Spectrum(1.0f), Spectrum(eta[0], eta[1], eta[2]), Spectrum(k[0], k[1], k[2])));
}
float ref_roughness_to_alpha(float r) { return GGX::roughnessToAlpha(r); }

// Metal preset eta/k RGB (material/metal.cpp:74-140).
int ref_metal_preset(const char *name, float *eta_k) {
  float p[8] = {0, 0, 0, 0, 0, 0, 0.1f, 0};
  auto m = std::dynamic_pointer_cast<Metal>(make_material(1, p, name));
  if (!m) return -1;
  Intersection is(Point3f(0, 0, 0));
  put3(eta_k, m->eta->evaluate(is));
  put3(eta_k + 3, m->k->evaluate(is));
  return 0;
}

// BSDF::f and BSDF::pdf.  frame = p ng ns u v (15 floats); params as make_material.
void ref_bsdf_f_pdf(int kind, const float *params, const char *metal_name, int photon,
                    const float *frame, const float *wi, const float *wo, int type,
                    float *f, float *pdf) {
  Intersection is = make_isect(frame);
  auto mat = make_material(kind, params, metal_name);
  auto bsdf = mat->getBSDF(is, photon ? TransportMode::Photon : TransportMode::Path);
  put3(f, bsdf->f(v3(wi), v3(wo), BxDFType(type)));
  *pdf = bsdf->pdf(v3(wi), v3(wo), BxDFType(type));
}

void ref_bsdf_sample(int kind, const float *params, const char *metal_name, int photon,
                     const float *frame, const float *wo, const float *u, int type,
                     float *f, float *wi, float *pdf, int *stype) {
  Intersection is = make_isect(frame);
  auto mat = make_material(kind, params, metal_name);
  auto bsdf = mat->getBSDF(is, photon ? TransportMode::Photon : TransportMode::Path);
  Vector3f w(0, 0, 0);
  Float p = -1;
  BxDFType t = BxDFType(-1);
  put3(f, bsdf->sample_f(&w, v3(wo), Point2f(u[0], u[1]), &p, BxDFType(type), &t));
  put3(wi, w);
  *pdf = p;
  *stype = (int)t;
}

// Camera: lookat (9), fov, W, H.  Returns the camera-to-world matrix (16,
// row-major), tan_half_fov, aspect and the ray for pixel (i, j), index.
void ref_camera(const float *lookat, float fov, int W, int H, int i, int j, long long index,
                float *mat16, float *tan_aspect, float *ray6) {
  Transform t = lookAt(Point3f(lookat[0], lookat[1], lookat[2]),
                       Point3f(lookat[3], lookat[4], lookat[5]),
                       Vector3f(lookat[6], lookat[7], lookat[8]));
  auto film = std::make_shared<RenderTarget>(W, H);
  Pinhole cam(t, fov, film);
  const Matrix4f &m = t.getMatrix();
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) mat16[4 * r + c] = m.m[r][c];
  tan_aspect[0] = cam.tan_half_fov;
  tan_aspect[1] = cam.aspectRatio;
  SobolSampler base(0);
  auto s = base.clone(index);
  Ray ray = cam.generateRay(Point2f((Float)i, (Float)j), *s);
  put3(ray6, ray.o);
  put3(ray6 + 3, ray.d);
}

// createTransform over (key, value) pairs (transform.cpp:218-252); out 32 = m, mInv.
void ref_create_transform(int n, const char **keys, const char **values, float *out) {
  ParamVec pv;
  for (int i = 0; i < n; ++i) pv.addParam(keys[i], values[i]);
  Transform t = createTransform(pv);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      out[4 * r + c] = t.getMatrix().m[r][c];
      out[16 + 4 * r + c] = t.getInverseMatrix().m[r][c];
    }
}

// Point (type 1) or spot (type 2) light built as the loader builds it
// (point.cpp:34-40, spot.cpp:46-59).  For the spot light the object is
// constructed in zeroed storage so the self-initialised cosFalloffStart
// (spot.h:25) reads 0 degrees.  info: power(3), cosTotal, cosFalloffStart,
// w2l(9), l2w(9), pos(3).
struct RefLightBuf { alignas(64) unsigned char bytes[sizeof(SpotLight) > sizeof(PointLight) ? sizeof(SpotLight) : sizeof(PointLight)]; };

static Light *build_light(int type, const float *I, const float *from, const float *to, float angle,
                          float delta, RefLightBuf *buf) {
  std::memset(buf, 0, sizeof(*buf));
  if (type == 1) {
    Transform l2w = translate(Vector3f(from[0], from[1], from[2]));
    return new (buf->bytes) PointLight(l2w, Spectrum(I[0], I[1], I[2]));
  }
  Point3f pf(from[0], from[1], from[2]), pt(to[0], to[1], to[2]);
  Vector3f dir = normalize(pt - pf);
  Vector3f x, y;
  coordinateSystem(dir, &x, &y);
  Transform dirToZ(Matrix4f(x.x, x.y, x.z, 0.0f, y.x, y.y, y.z, 0.0f, dir.x, dir.y, dir.z, 0.0f,
                            0.0f, 0.0f, 0.0f, 1.0f));
  Transform l2w = translate(Vector3f(pf)) * inverse(dirToZ);
  return new (buf->bytes) SpotLight(l2w, Spectrum(I[0], I[1], I[2]), angle, angle - delta);
}

void ref_light_info(int type, const float *I, const float *from, const float *to, float angle,
                    float delta, float *info) {
  RefLightBuf buf;
  Light *l = build_light(type, I, from, to, angle, delta, &buf);
  put3(info, l->power());
  const Matrix4f &w2l = l->world2local.getMatrix();
  const Matrix4f &l2w = l->local2world.getMatrix();
  if (type == 2) {
    SpotLight *s = static_cast<SpotLight *>(l);
    info[3] = s->cosTotalWidth;
    info[4] = s->cosFalloffStart;
    put3(info + 23, s->p);
  } else {
    info[3] = info[4] = 0;
    put3(info + 23, static_cast<PointLight *>(l)->p);
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      info[5 + 3 * r + c] = w2l.m[r][c];
      info[14 + 3 * r + c] = l2w.m[r][c];
    }
  l->~Light();
}

// sample_Li at (p, ng): out Li(3) wi(3) pdf  shadow o(3) d(3) tmax
void ref_light_sample_li(int type, const float *I, const float *from, const float *to,
                         float angle, float delta, const float *p, const float *ng,
                         const float *u, float *out) {
  RefLightBuf buf;
  Light *l = build_light(type, I, from, to, angle, delta, &buf);
  Intersection is(Point3f(p[0], p[1], p[2]));
  is.ng = Normal3f(ng[0], ng[1], ng[2]);
  Vector3f wi;
  Float pdf = 0;
  ShadowTester sdt;
  Spectrum Li = l->sample_Li(is, Point2f(u[0], u[1]), &wi, &pdf, &sdt);
  put3(out, Li);
  put3(out + 3, wi);
  out[6] = pdf;
  Ray r = sdt.p0.spawnRay(sdt.p1);
  put3(out + 7, r.o);
  put3(out + 10, r.d);
  out[13] = r.tMax;
  l->~Light();
}

// sample_Le: out Le(3) o(3) d(3) n(3) pdfPos pdfDir
void ref_light_sample_le(int type, const float *I, const float *from, const float *to,
                         float angle, float delta, const float *u1, const float *u2, float *out) {
  RefLightBuf buf;
  Light *l = build_light(type, I, from, to, angle, delta, &buf);
  Ray r;
  Normal3f n;
  Float pp = 0, pd = 0;
  Spectrum Le = l->sample_Le(Point2f(u1[0], u1[1]), Point2f(u2[0], u2[1]), &r, &n, &pp, &pd);
  put3(out, Le);
  put3(out + 3, r.o);
  put3(out + 6, r.d);
  put3(out + 9, n);
  out[12] = pp;
  out[13] = pd;
  l->~Light();
}

// Distribution1D (sampling.h:71-125) over func[0..n) at u: [0] sampleDiscrete
// offset, [1] its pdf, [2] sampleContinuous value, [3] its pdf, [4] its offset
void ref_distribution1d(int n, const float *func, float u, float *out) {
  Distribution1D d(func, n);
  Float pdf = 0;
  out[0] = (float)d.sampleDiscrete(u, &pdf);
  out[1] = pdf;
  int off = 0;
  out[2] = d.sampleContinuous(u, &pdf, &off);
  out[3] = pdf;
  out[4] = (float)off;
}

// offsetRayOrigin + spawnRay(d) (geometry.h:850-863, intersection.h:46-49)
void ref_spawn(const float *p, const float *ng, const float *d, float *o3) {
  Intersection is(Point3f(p[0], p[1], p[2]));
  is.ng = Normal3f(ng[0], ng[1], ng[2]);
  put3(o3, is.spawnRay(Vector3f(d[0], d[1], d[2])).o);
}


// ------------------------------------------------------------------ infinite lights
// The Scene a light's preprocess reads cannot be built (Embree): the world
// sphere is set directly, and pdf_Li's unused Scene argument is a stand-in
// that is never touched.
namespace {
const Scene &no_scene() {
  alignas(Scene) static char storage[sizeof(Scene)];
  return *reinterpret_cast<const Scene *>(storage);
}
Transform xform16(const float *m, const float *mi) {
  Matrix4f a, b;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) { a.m[r][c] = m[4 * r + c]; b.m[r][c] = mi[4 * r + c]; }
  return Transform(a, b);
}
void put_ray(float *o, const Ray &r) {
  put3(o, r.o); put3(o + 3, r.d); o[6] = r.tMax;
}
// out: Li(3) wi(3) pdf shadow ray o(3) d(3) tmax
void sample_li(const Light &l, const float *p, const float *ng, const float *u, float *out) {
  Intersection is(Point3f(p[0], p[1], p[2]));
  is.ng = Normal3f(ng[0], ng[1], ng[2]);
  Vector3f wi(0, 0, 0);
  Float pdf = 0;
  ShadowTester sdt;
  Spectrum Li = l.sample_Li(is, Point2f(u[0], u[1]), &wi, &pdf, &sdt);
  put3(out, Li); put3(out + 3, wi); out[6] = pdf;
  if (pdf > 0 && !Li.isBlack()) put_ray(out + 7, sdt.p0.spawnRay(sdt.p1));
}
// out: Le(3) ray o(3) d(3) nLight(3) pdfPos pdfDir
void sample_le(const Light &l, const float *u1, const float *u2, float *out) {
  Ray r(Point3f(0, 0, 0), Vector3f(0, 0, 0));
  Normal3f n(0, 0, 0);
  Float pp = 0, pd = 0;
  Spectrum Le = l.sample_Le(Point2f(u1[0], u1[1]), Point2f(u2[0], u2[1]), &r, &n, &pp, &pd);
  put3(out, Le); put3(out + 3, r.o); put3(out + 6, r.d); put3(out + 9, n); out[12] = pp; out[13] = pd;
}
}  // namespace

// loadImage (texture/imagetexture.cpp:31-49): returns channels, fills rgb (w*h*3)
int ref_load_image(const char *path, int *w, int *h, float *rgb, int cap) {
  int ww, hh, ch;
  unsigned char *d = stbi_load(path, &ww, &hh, &ch, 0);
  if (!d) return -1;
  stbi_image_free(d);
  auto im = loadImage(path);
  *w = im->width; *h = im->height;
  for (int i = 0; i < im->width * im->height && 3 * i + 2 < cap; ++i) put3(rgb + 3 * i, im->data[i]);
  return ch;
}

// Envmap(light2world, L, 1, path) over an image file (or "" = constant L)
void *ref_envmap_new(const char *path, const float *L, const float *m, const float *mi, const float *center,
                     float radius) {
  auto *e = new Envmap(xform16(m, mi), Spectrum(L[0], L[1], L[2]), 1, path ? path : "");
  e->worldCenter = Point3f(center[0], center[1], center[2]);
  e->worldRadius = radius;
  return e;
}
void ref_envmap_free(void *e) { delete (Envmap *)e; }
void ref_envmap_le(void *e, const float *d, float *out) {
  put3(out, ((Envmap *)e)->Le(Ray(Point3f(0, 0, 0), v3(d))));
}
float ref_envmap_pdf_li(void *e, const float *wi) {
  Intersection is(Point3f(0, 0, 0));
  return ((Envmap *)e)->pdf_Li(is, v3(wi), no_scene());
}
void ref_envmap_sample_li(void *e, const float *p, const float *ng, const float *u, float *out) {
  sample_li(*(Envmap *)e, p, ng, u, out);
}
void ref_envmap_sample_le(void *e, const float *u1, const float *u2, float *out) {
  sample_le(*(Envmap *)e, u1, u2, out);
}
void ref_envmap_power(void *e, float *out) { put3(out, ((Envmap *)e)->power()); }

// createDistantLight (distant.cpp:58-65) from L / from / to strings
void *ref_distant_new(const char *L, const char *from, const char *to, const float *center, float radius) {
  ParamSet ps;
  ps.setParam("L", L);
  ps.setParam("from", from);
  ps.setParam("to", to);
  auto d = createDistantLight(ps);
  auto *l = new DistantLight(*d);
  l->worldCenter = Point3f(center[0], center[1], center[2]);
  l->worldRadius = radius;
  return l;
}
void ref_distant_free(void *l) { delete (DistantLight *)l; }
void ref_distant_dir(void *l, float *out) { put3(out, ((DistantLight *)l)->dir); }
void ref_distant_sample_li(void *l, const float *p, const float *ng, const float *u, float *out) {
  sample_li(*(DistantLight *)l, p, ng, u, out);
}
void ref_distant_sample_le(void *l, const float *u1, const float *u2, float *out) {
  sample_le(*(DistantLight *)l, u1, u2, out);
}
void ref_distant_power(void *l, float *out) { put3(out, ((DistantLight *)l)->power()); }

// RenderTarget with the given sums and counts -> getAveraged (avg, W*H*3)
// -> saveAsBmp(path) (texture/rendertarget.cpp:28-37, core/texture.cpp:77-137)
int ref_film_average_bmp(int W, int H, const float *sum, const int *count, float *avg, const char *path) {
  RenderTarget rt(W, H);
  for (int i = 0; i < W * H; ++i) {
    rt.data[i] = Spectrum(sum[3 * i], sum[3 * i + 1], sum[3 * i + 2]);
    rt.counter[i] = count[i];
  }
  auto a = rt.getAveraged();
  for (int i = 0; i < W * H; ++i) put3(avg + 3 * i, a->data[i]);
  return path ? (saveAsBmp(a.get(), path) ? 0 : -1) : 0;
}

// createEnvMap's light-to-world / world-to-light (envmap.cpp:108-128) for a
// <Transform> parameter list and a worldUp string (constant map: no image)
void ref_envmap_xform(int n, const char **keys, const char **values, const char *world_up, float *out) {
  ParamVec pv;
  for (int i = 0; i < n; ++i) pv.addParam(keys[i], values[i]);
  Transform t = createTransform(pv);
  ParamSet ps;
  ps.setParam("worldUp", world_up);
  ps.setParam("nLightSamples", "1");
  auto e = createEnvMap(t, ps);
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      out[4 * r + c] = e->local2world.getMatrix().m[r][c];
      out[16 + 4 * r + c] = e->world2local.getMatrix().m[r][c];
    }
}

}  // extern "C"
