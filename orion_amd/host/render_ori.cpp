// Renders an .ori scene through the C++ host layer, as orion's main.cpp does
// (parse -> integrator->render(*scene) -> getFilm()->getAveraged() ->
// saveAsBmp), writing the averaged image as BMP, or as FP32 PFM / EXR, by the
// output file's extension.
//
//   orion_render scene.ori out.{bmp,pfm,exr} [device]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "orion_gpu.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s scene.ori out.{bmp,pfm,exr} [device]\n", argv[0]);
    return 2;
  }
  orion_ctx ctx = nullptr;
  if (!orion_gpu::check(orion_ctx_create(argc > 3 ? std::atoi(argv[3]) : 0, &ctx), "ctx")) return 1;
  int rc = 0;
  {
    auto t0 = std::chrono::steady_clock::now();
    orion_gpu::Parser parser(argv[1], ctx);
    const auto &opt = parser.getRenderOption();
    if (!opt || !opt->scene->ok()) {
      rc = 1;
    } else {
      auto t1 = std::chrono::steady_clock::now();
      opt->integrator->render(*opt->scene);
      auto t2 = std::chrono::steady_clock::now();
      auto film = opt->camera->getFilm();
      std::vector<float> img = film->getAveraged();
      std::string out = argv[2];
      std::string ext = out.size() > 4 ? out.substr(out.size() - 4) : "";
      int fmt = ext == ".pfm" ? ORION_IMAGE_PFM : (ext == ".exr" ? ORION_IMAGE_EXR : ORION_IMAGE_BMP);
      if (!orion_gpu::check(orion_film_write(out.c_str(), fmt, film->getWidth(), film->getHeight(), img.data()),
                            "write image"))
        rc = 1;
      const auto &st = opt->integrator->stats;
      std::printf("{\"load_ms\": %.1f, \"render_ms\": %.2f, \"device_ms\": %.2f, \"paths\": %lld, "
                  "\"closest_rays\": %lld, \"shadow_rays\": %lld, \"photons\": %lld}\n",
                  std::chrono::duration<double, std::milli>(t1 - t0).count(),
                  std::chrono::duration<double, std::milli>(t2 - t1).count(), st.ms, (long long)st.paths,
                  (long long)st.closest_rays, (long long)st.shadow_rays, (long long)st.photons);
    }
  }
  orion_ctx_destroy(ctx);
  return rc;
}
