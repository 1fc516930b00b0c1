// C++ host layer over the C ABI (include/orion_amd.h), shaped like orion's
// own interfaces so a caller of orion finds the same objects:
//
//   orion_gpu::Parser          Parser (util/parser.h:44-65): .ori -> RenderOption
//   orion_gpu::RenderOption    RenderOption (util/parser.h:44-53)
//   orion_gpu::Scene           Scene (core/scene.h:16-28) backed by a GpuKernel
//   orion_gpu::GpuKernel       Kernel (core/kernel.h:16-25): clear / build /
//                              intersect / worldBound, plus a batched form
//   orion_gpu::Camera          Camera (core/camera.h:17-49) with its film
//   orion_gpu::RenderTarget    RenderTarget (texture/rendertarget.h): sum + count
//   orion_gpu::SobolSampler    Sampler (core/sampler.h:15-34) for host-side use
//   orion_gpu::Integrator      Integrator (core/integrator.h:17-33)
//   orion_gpu::GpuPathTracing  PathTracing (integrator/path.h)
//   orion_gpu::GpuSPPM         SPPM (integrator/sppm.h)
//
// Error behaviour follows orion: failures print a message (CHECK_INFO style,
// orion.h:33-37) and the object stays usable but empty; ok() reports it.
#pragma once
#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"

namespace orion_gpu {

inline bool check(int rc, const char *what) {
  if (rc == ORION_OK) return true;
  std::printf("%s failed (%d): %s\n\t|file: %s\n", what, rc, orion_last_error(), __FILE__);
  return false;
}

struct Ray {  // core/geometry.h:596-610
  float o[3], d[3];
  float t = 0.f, tMax = 3.40282347e+38f;
};
struct Intersection {  // the fields GpuKernel::intersect fills (core/kernel.cpp:127-172 inputs)
  int primID = -1;
  float t = 0, u = 0, v = 0;
};

class RenderTarget {  // texture/rendertarget.h: per-pixel sum + count
 public:
  RenderTarget(int w, int h) : width(w), height(h), data(3 * (size_t)w * h, 0.f), counter((size_t)w * h, 0) {}
  int getWidth() const { return width; }
  int getHeight() const { return height; }
  // getAveraged (rendertarget.cpp:28-37): data / counter
  std::vector<float> getAveraged() const {
    std::vector<float> r(data.size());
    for (size_t i = 0; i < counter.size(); ++i)
      for (int c = 0; c < 3; ++c) r[3 * i + c] = data[3 * i + c] / counter[i];
    return r;
  }
  int width, height;
  std::vector<float> data;
  std::vector<int> counter;
};

class Camera {
 public:
  explicit Camera(const float record[16]) {
    for (int i = 0; i < 16; ++i) rec[i] = record[i];
    film = std::make_shared<RenderTarget>((int)rec[14], (int)rec[15]);
  }
  std::shared_ptr<RenderTarget> getFilm() const { return film; }
  float rec[16];

 private:
  std::shared_ptr<RenderTarget> film;
};

class SobolSampler {  // host-side Sampler with orion's dimension bookkeeping
 public:
  explicit SobolSampler(int64_t index) : index(index), dimention(5) {}
  int64_t index;
  int dimention;
};

// Kernel replacement: the device BVH.  build() uploads; intersect() is the
// batched closest-hit (per-ray calls are supported for parity but cost a
// round trip each -- orion's per-ray interface does not suit a GPU).
class GpuKernel {
 public:
  GpuKernel(orion_ctx ctx, orion_hscene hs) : ctx_(ctx), hs_(hs) {}
  ~GpuKernel() { clear(); }
  void clear() {
    if (scene_) orion_scene_destroy(scene_);
    scene_ = nullptr;
  }
  bool build() {
    clear();
    return check(orion_scene_upload(ctx_, hs_, &scene_), "GpuKernel::build");
  }
  bool intersect(const Ray &ray, Intersection *isec) const {
    float r[8] = {ray.o[0], ray.o[1], ray.o[2], ray.d[0], ray.d[1], ray.d[2], ray.t, ray.tMax};
    int tie = 0;
    if (!check(orion_intersect_closest(scene_, 1, r, &isec->t, &isec->primID, &isec->u, &isec->v, &tie, 0),
               "GpuKernel::intersect"))
      return false;
    return isec->primID >= 0;
  }
  // batched form: rays SoA [8][n] (host), hits per ray
  bool intersect(int n, const float *rays_soa, std::vector<Intersection> *out) const {
    std::vector<float> t(n), u(n), v(n);
    std::vector<int> prim(n), tie(n);
    if (!check(orion_intersect_closest(scene_, n, rays_soa, t.data(), prim.data(), u.data(), v.data(), tie.data(), 0),
               "GpuKernel::intersect"))
      return false;
    out->resize(n);
    for (int i = 0; i < n; ++i) (*out)[i] = Intersection{prim[i], t[i], u[i], v[i]};
    return true;
  }
  orion_scene device_scene() const { return scene_; }

 private:
  orion_ctx ctx_;
  orion_hscene hs_;
  orion_scene scene_ = nullptr;
};

class Scene {
 public:
  Scene(orion_ctx ctx, orion_hscene hs) : kernel(new GpuKernel(ctx, hs)) { ok_ = kernel->build(); }
  bool intersect(const Ray &ray, Intersection *isec) const { return kernel->intersect(ray, isec); }
  bool ok() const { return ok_; }
  std::unique_ptr<GpuKernel> kernel;

 private:
  bool ok_ = false;
};

class Integrator {  // core/integrator.h:17-33
 public:
  Integrator(std::shared_ptr<Camera> camera, std::shared_ptr<SobolSampler> sampler)
      : camera(std::move(camera)), sampler(std::move(sampler)) {}
  virtual ~Integrator() {}
  virtual void render(const Scene &scene) = 0;
  orion_render_stats stats{};

 protected:
  std::shared_ptr<Camera> camera;
  std::shared_ptr<SobolSampler> sampler;
  bool read_film(const Scene &scene) {
    auto film = camera->getFilm();
    return check(orion_film_read(scene.kernel->device_scene(), film->data.data(), film->counter.data()),
                 "film read");
  }
};

class GpuPathTracing : public Integrator {  // integrator/path.cpp:15-37
 public:
  GpuPathTracing(std::shared_ptr<Camera> c, std::shared_ptr<SobolSampler> s, int maxDepth, int nSamples)
      : Integrator(std::move(c), std::move(s)), maxDepth(maxDepth), nSamples(nSamples) {}
  void render(const Scene &scene) override {
    orion_scene ds = scene.kernel->device_scene();
    if (!check(orion_film_clear(ds), "film clear")) return;
    if (!check(orion_render_path(ds, 0, nSamples, maxDepth, nullptr, &stats), "PathTracing::render")) return;
    read_film(scene);
  }
  int maxDepth, nSamples;
};

class GpuSPPM : public Integrator {  // integrator/sppm.cpp:58-322
 public:
  GpuSPPM(std::shared_ptr<Camera> c, std::shared_ptr<SobolSampler> s, int maxDepth, int nIterations,
          int photonsPerIteration, float initialSearchRadius, int hashSize = 0)
      : Integrator(std::move(c), std::move(s)), maxDepth(maxDepth), nIterations(nIterations),
        photonPerIteration(photonsPerIteration), initialSearchRadius(initialSearchRadius), hashSize(hashSize) {}
  void render(const Scene &scene) override {
    orion_scene ds = scene.kernel->device_scene();
    if (!check(orion_film_clear(ds), "film clear")) return;
    if (!check(orion_render_sppm(ds, nIterations, photonPerIteration, initialSearchRadius, maxDepth, hashSize,
                                 nullptr, &stats),
               "SPPM::render"))
      return;
    read_film(scene);
  }
  int maxDepth, nIterations, photonPerIteration;
  float initialSearchRadius;
  int hashSize;
};

struct RenderOption {  // util/parser.h:44-53
  std::shared_ptr<Scene> scene;
  std::shared_ptr<Camera> camera;
  std::shared_ptr<Integrator> integrator;
  std::shared_ptr<SobolSampler> sampler;
  int nSamples = 1;
};

class Parser {  // util/parser.h:55-65: Parser(path) -> getRenderOption()
 public:
  Parser(const std::string &path, orion_ctx ctx) {
    if (!check(orion_hscene_load_ori(path.c_str(), &hs_), "Parser")) return;
    orion_scene_info info;
    orion_hscene_info(hs_, &info);
    const float *cam = nullptr;
    orion_hscene_records(hs_, nullptr, nullptr, nullptr, nullptr, nullptr, &cam);
    opt_.reset(new RenderOption());
    opt_->camera = std::make_shared<Camera>(cam);
    opt_->sampler = std::make_shared<SobolSampler>(3);  // createSobolSampler (sobol.h)
    if (info.integrator == 0) {
      opt_->integrator = std::make_shared<GpuPathTracing>(opt_->camera, opt_->sampler, info.max_depth, info.spp);
      opt_->nSamples = info.spp;
    } else {
      opt_->integrator = std::make_shared<GpuSPPM>(opt_->camera, opt_->sampler, info.max_depth,
                                                   info.sppm_iterations, info.sppm_photons, info.sppm_radius);
    }
    opt_->scene = std::make_shared<Scene>(ctx, hs_);
  }
  ~Parser() {
    if (opt_) opt_->scene.reset();
    if (hs_) orion_hscene_destroy(hs_);
  }
  const std::unique_ptr<RenderOption> &getRenderOption() const { return opt_; }

 private:
  orion_hscene hs_ = nullptr;
  std::unique_ptr<RenderOption> opt_;
};

}  // namespace orion_gpu
