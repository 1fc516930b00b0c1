// orion drop-in (see orion_dropin.h): compiled in orion's tree against its
// headers, linked with liborion_amd.so.
#include "orion_dropin.h"

#include <core/camera.h>
#include <core/intersection.h>
#include <core/primitive.h>
#include <core/spectrum.h>
#include <core/triangle.h>
#include <sampler/sobol.h>
#include <texture/rendertarget.h>
#include <util/param.h>
#include <util/strutil.h>

#include <iostream>
#include <vector>

ORION_NAMESPACE_BEGIN

namespace {

// A failed library call is printed in orion's way (CHECK_INFO) and the
// caller carries on without the result.
bool succeeded(int rc, const char *call) {
  CHECK_INFO(rc == ORION_OK, std::string("orion_amd: ") + call + ": " + orion_last_error());
  return rc == ORION_OK;
}

// A context, the .ori's resolved records and their device copy.
struct GpuRun {
  orion_ctx ctx = nullptr;
  orion_hscene hscene = nullptr;
  orion_scene scene = nullptr;
  int width = 0, height = 0;
  ~GpuRun() {
    if (scene) orion_scene_destroy(scene);
    if (hscene) orion_hscene_destroy(hscene);
    if (ctx) orion_ctx_destroy(ctx);
  }
  // Loads the .ori the Parser read (its records equal orion's parse) and
  // uploads it.  The GPU draws Sobol samples and renders into a film of the
  // camera's size; anything else is refused.
  bool open(int device, const std::string &path, const Sampler *sampler, const RenderTarget &film) {
    CHECK_INFO(dynamic_cast<const SobolSampler *>(sampler) != nullptr,
               "orion_amd: the GPU integrators draw Sobol samples: use <Sampler type=\"sobol\"/>");
    if (!dynamic_cast<const SobolSampler *>(sampler)) return false;
    if (!succeeded(orion_hscene_load_ori(path.c_str(), &hscene), "orion_hscene_load_ori")) return false;
    orion_scene_info info;
    if (!succeeded(orion_hscene_info(hscene, &info), "orion_hscene_info")) return false;
    width = info.width;
    height = info.height;
    CHECK_INFO(width == film.getWidth() && height == film.getHeight(),
               "orion_amd: the camera's film is not the .ori's film size");
    if (width != film.getWidth() || height != film.getHeight()) return false;
    if (!succeeded(orion_ctx_create(device, &ctx), "orion_ctx_create")) return false;
    return succeeded(orion_scene_upload(ctx, hscene, &scene), "orion_scene_upload");
  }
  // The device film into orion's RenderTarget through its own accumulate
  // (rendertarget.cpp:22-27): per pixel the sum of its samples, added in
  // sample order on the device, then zero for each further sample, so the
  // target holds orion's sum and count and getAveraged() divides as orion's.
  void accumulate_into(RenderTarget &film) const {
    std::vector<float> sum((size_t)width * height * 3);
    std::vector<int> count((size_t)width * height);
    if (!succeeded(orion_film_read(scene, sum.data(), count.data()), "orion_film_read")) return;
    for (int j = 0; j < height; ++j)
      for (int i = 0; i < width; ++i) {
        const size_t q = (size_t)j * width + i;
        film.acumulate(i, j, Spectrum(sum[3 * q], sum[3 * q + 1], sum[3 * q + 2]));
        for (int k = 1; k < count[q]; ++k) film.acumulate(i, j, Spectrum(0.f));
      }
  }
};

}  // namespace

// ---------------------------------------------------------------- GpuKernel
GpuKernel::GpuKernel(const std::vector<std::shared_ptr<Primitive>> &primitives, int device)
    : primitives(primitives), device(device) {}

GpuKernel::~GpuKernel() { clear(); }

void GpuKernel::clear() {
  if (scene) orion_scene_destroy(scene);
  if (hscene) orion_hscene_destroy(hscene);
  if (ctx) orion_ctx_destroy(ctx);
  scene = nullptr;
  hscene = nullptr;
  ctx = nullptr;
}

void GpuKernel::build() {
  clear();
  const int n = (int)primitives.size();
  // world-space corners and normals per primitive (Mesh is de-indexed and
  // pre-transformed, core/triangle.cpp:137-151)
  std::vector<float> pn((size_t)n * 18);
  bound = Bounds3f();
  for (int i = 0; i < n; ++i) {
    const Triangle &tri = *primitives[i]->triangle;
    float *r = &pn[(size_t)i * 18];
    for (int k = 0; k < 3; ++k) {
      const Point3f &p = tri.mesh->p[tri.triNumber * 3 + k];
      const Normal3f &nn = tri.mesh->n[tri.triNumber * 3 + k];
      r[3 * k] = p.x; r[3 * k + 1] = p.y; r[3 * k + 2] = p.z;
      r[9 + 3 * k] = nn.x; r[9 + 3 * k + 1] = nn.y; r[9 + 3 * k + 2] = nn.z;
    }
    bound = Union(bound, primitives[i]->worldBound());
  }
  // intersection only: one placeholder material, no light, a 1 x 1 view
  std::vector<int> mat((size_t)n, 0), light((size_t)n, -1);
  float matte[ORION_MAT_FLOATS] = {0.f, 0.5f, 0.5f, 0.5f};
  float no_light[32] = {0.f};
  float view[16] = {0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 1.f, 1.f};
  if (!succeeded(orion_ctx_create(device, &ctx), "orion_ctx_create")) return;
  if (!succeeded(orion_hscene_from_records(n, pn.data(), mat.data(), light.data(), 1, matte, 0, no_light, view,
                                           &hscene),
                 "orion_hscene_from_records"))
    return;
  succeeded(orion_scene_upload(ctx, hscene, &scene), "orion_scene_upload");
}

bool GpuKernel::intersect(const Ray &ray, Intersection *isec) const {
  bool hit = false;
  intersect(1, &ray, isec, &hit);
  return hit;
}

void GpuKernel::intersect(int n, const Ray *rays, Intersection *isecs, bool *hit) const {
  for (int i = 0; i < n; ++i) hit[i] = false;
  if (!scene || n <= 0) return;
  // rays as the C ABI takes them: [8][n] = o, d, [t, tMax] (rtcore's
  // [tnear, tfar], kernel.cpp:113-118)
  std::vector<float> soa((size_t)8 * n), t((size_t)n), u((size_t)n), v((size_t)n);
  std::vector<int> prim((size_t)n);
  for (int i = 0; i < n; ++i) {
    const Ray &r = rays[i];
    const float f[8] = {r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z, r.t, r.tMax};
    for (int a = 0; a < 8; ++a) soa[(size_t)a * n + i] = f[a];
  }
  if (!succeeded(orion_intersect_closest(scene, n, soa.data(), t.data(), prim.data(), u.data(), v.data(), nullptr, 0),
                 "orion_intersect_closest"))
    return;
  for (int i = 0; i < n; ++i)
    if (prim[i] >= 0) {
      hit[i] = true;
      rebuild(rays[i], prim[i], u[i], v[i], t[i], &isecs[i]);
    }
}

// EmbreeKernel::intersect's hit record (kernel.cpp:127-172): point, uv and
// shading normal from the barycentrics (the shading normal not normalized),
// geometric normal of the corners, the ray's side of the shading normal,
// the shading frame u, v around ns, and the primitive.
void GpuKernel::rebuild(const Ray &ray, int primID, float bu, float bv, float t, Intersection *isec) const {
  const Primitive &prim = *primitives[primID];
  const Triangle &tri = *prim.triangle;
  const Mesh &m = *tri.mesh;
  const int c = tri.triNumber * 3;
  const Float w = 1 - bu - bv;
  const Point3f p = w * m.p[c] + bu * m.p[c + 1] + bv * m.p[c + 2];
  const Point2f uv = w * m.uv[c] + bu * m.uv[c + 1] + bv * m.uv[c + 2];
  const Normal3f ns = m.n[c] * w + m.n[c + 1] * bu + m.n[c + 2] * bv;
  const Normal3f ng = Normal3f(normalize(cross(m.p[c + 1] - m.p[c], m.p[c + 2] - m.p[c])));
  const bool front = dot(ray.o - m.p[c], ns) > 0;
  *isec = Intersection(p, ng, ns, uv, t, front);
  const Vector3f e = normalize(m.p[c + 1] - m.p[c]);
  const Vector3f fv = normalize(cross(ns, e));
  isec->u = normalize(cross(fv, ns));
  isec->v = fv;
  isec->primitive = &prim;
}

Bounds3f GpuKernel::worldBound() const { return bound; }

// ---------------------------------------------------------------- integrators
GpuPathTracing::GpuPathTracing(const std::shared_ptr<Camera> &camera, const std::shared_ptr<Sampler> &sampler,
                               int maxDepth, int nSamples, const std::string &scenePath, int device)
    : Integrator(camera, sampler), maxDepth(maxDepth), nSamples(nSamples), device(device), scenePath(scenePath) {}

void GpuPathTracing::render(const Scene &) {
  RenderTarget &film = *camera->getFilm();
  GpuRun run;
  if (!run.open(device, scenePath, sampler.get(), film)) return;
  orion_render_stats st;
  if (!succeeded(orion_render_path(run.scene, 0, nSamples, maxDepth, nullptr, &st), "orion_render_path")) return;
  run.accumulate_into(film);
}

GpuSPPM::GpuSPPM(const std::shared_ptr<Camera> &camera, const std::shared_ptr<Sampler> &sampler, int nIterations,
                 int photonsPerIteration, int maxDepth, Float initialSearchRadius, const std::string &scenePath,
                 int device)
    : Integrator(camera, sampler),
      nIterations(nIterations),
      photonsPerIteration(photonsPerIteration),
      maxDepth(maxDepth),
      device(device),
      initialSearchRadius(initialSearchRadius),
      scenePath(scenePath) {}

void GpuSPPM::render(const Scene &) {
  RenderTarget &film = *camera->getFilm();
  GpuRun run;
  if (!run.open(device, scenePath, sampler.get(), film)) return;
  orion_render_stats st;
  // hash_size 0: nPixels cells, as orion sizes its grid (sppm.cpp:147)
  if (!succeeded(orion_render_sppm(run.scene, nIterations, photonsPerIteration, (float)initialSearchRadius, maxDepth,
                                   0, nullptr, &st),
                 "orion_render_sppm"))
    return;
  run.accumulate_into(film);
}

std::shared_ptr<GpuPathTracing> createGpuPathTracingIntegrator(const std::shared_ptr<Camera> &camera,
                                                               const std::shared_ptr<Sampler> &sampler,
                                                               const ParamSet &param, const std::string &scenePath) {
  const int maxDepth = parseInt(param.getParam("maxDepth"));
  const int nSamples = parseInt(param.getParam("nSamplesPerPixel"));
  return std::make_shared<GpuPathTracing>(camera, sampler, maxDepth, nSamples, scenePath);
}

std::shared_ptr<GpuSPPM> createGpuSPPMIntegrator(const std::shared_ptr<Camera> &camera,
                                                 const std::shared_ptr<Sampler> &sampler, const ParamSet &param,
                                                 const std::string &scenePath) {
  const int nIterations = parseInt(param.getParam("numIterations"));
  const int maxDepth = parseInt(param.getParam("maxDepth"));
  int photonsPerIter = parseInt(param.getParam("photonsPerIteration"));
  if (photonsPerIter < 0) photonsPerIter = camera->getFilm()->getHeight() * camera->getFilm()->getWidth();
  const Float radius = parseFloat(param.getParam("initRadius"));
  return std::make_shared<GpuSPPM>(camera, sampler, nIterations, photonsPerIter, maxDepth, radius, scenePath);
}

ORION_NAMESPACE_END
