// orion drop-in: the GPU path behind orion's own Kernel and Integrator
// interfaces.  Compiled inside orion's source tree, against its headers, and
// linked with liborion_amd.so (include/orion_amd.h); three lines of orion
// change to use it:
//
//   core/scene.cpp:9      kernel = std::unique_ptr<Kernel>(new GpuKernel(primitives));
//   util/parser.cpp:130   renderOption->integrator = createGpuPathTracingIntegrator(
//                             renderOption->camera, renderOption->sampler, ps, doc->Value());
//   util/parser.cpp:132   renderOption->integrator = createGpuSPPMIntegrator(
//                             renderOption->camera, renderOption->sampler, ps, doc->Value());
//
// (doc->Value() is the .ori path the Parser opened.)  orion's material
// classes keep their textures private, so the integrators take the scene's
// resolved records from orion_hscene_load_ori on the same .ori file -- the
// loader whose records equal orion's parse (tests/test_host.py) -- and write
// their film into the camera's RenderTarget with orion's own accumulate, so
// getAveraged() and saveAsBmp() work unchanged.  Errors follow orion's
// convention: printed with CHECK_INFO, execution continues.
#pragma once
#ifndef ORION_AMD_DROPIN_H_
#define ORION_AMD_DROPIN_H_

#include <orion.h>
#include <core/geometry.h>
#include <core/integrator.h>
#include <core/kernel.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"

ORION_NAMESPACE_BEGIN

// Kernel on the GPU: replaces EmbreeKernel (core/kernel.h:27-42,
// core/kernel.cpp:63-174).  build() uploads the primitives' world-space
// triangles in Scene order (primID = index into `primitives`, as Embree's
// one mesh numbers them) and builds the BVH on the GPU side;
// intersect() casts one ray through orion_intersect_closest and rebuilds
// orion's Intersection exactly as EmbreeKernel::intersect does.  The GPU
// answers many rays at once: intersect(n, ...) casts a batch in one call.
class GpuKernel : public Kernel {
 public:
  explicit GpuKernel(const std::vector<std::shared_ptr<Primitive>> &primitives, int device = 0);
  ~GpuKernel() override;
  void clear() override;
  void build() override;
  bool intersect(const Ray &ray, Intersection *isec) const override;
  Bounds3f worldBound() const override;
  // closest hits of n rays; hit[i] false where ray i misses
  void intersect(int n, const Ray *rays, Intersection *isecs, bool *hit) const;

 private:
  void rebuild(const Ray &ray, int prim, float bu, float bv, float t, Intersection *isec) const;
  std::vector<std::shared_ptr<Primitive>> primitives;
  Bounds3f bound;
  int device;
  orion_ctx ctx = nullptr;
  orion_hscene hscene = nullptr;
  orion_scene scene = nullptr;
};

// PathTracing::render on the GPU (integrator/path.cpp:15-37): the
// configuration's samples [0, nSamples) of every pixel, Sobol indices
// k * W * H + j * W + i as orion draws them.
class GpuPathTracing : public Integrator {
 public:
  GpuPathTracing(const std::shared_ptr<Camera> &camera, const std::shared_ptr<Sampler> &sampler, int maxDepth,
                 int nSamples, const std::string &scenePath, int device = 0);
  void render(const Scene &scene) override;

 private:
  int maxDepth, nSamples, device;
  std::string scenePath;
};

// SPPM::render on the GPU (integrator/sppm.cpp:58-322), hash table of
// nPixels cells as orion sizes it.
class GpuSPPM : public Integrator {
 public:
  GpuSPPM(const std::shared_ptr<Camera> &camera, const std::shared_ptr<Sampler> &sampler, int nIterations,
          int photonsPerIteration, int maxDepth, Float initialSearchRadius, const std::string &scenePath,
          int device = 0);
  void render(const Scene &scene) override;

 private:
  int nIterations, photonsPerIteration, maxDepth, device;
  Float initialSearchRadius;
  std::string scenePath;
};

// Created like createPathTracingIntegrator / createSPPMIntegrator
// (path.cpp:108-113, sppm.cpp:324-333) from the <Integrator> ParamSet, plus
// the path of the .ori file being parsed.
std::shared_ptr<GpuPathTracing> createGpuPathTracingIntegrator(const std::shared_ptr<Camera> &camera,
                                                               const std::shared_ptr<Sampler> &sampler,
                                                               const ParamSet &param, const std::string &scenePath);
std::shared_ptr<GpuSPPM> createGpuSPPMIntegrator(const std::shared_ptr<Camera> &camera,
                                                 const std::shared_ptr<Sampler> &sampler, const ParamSet &param,
                                                 const std::string &scenePath);

ORION_NAMESPACE_END

#endif  // ORION_AMD_DROPIN_H_
