// Renders an .ori scene through the C++ host layer, as orion's main.cpp does
// (parse -> integrator->render(*scene) -> getFilm()->getAveraged()), and
// writes the averaged FP32 image as a PFM.
//
//   orion_render scene.ori out.pfm [device]
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "orion_gpu.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s scene.ori out.pfm [device]\n", argv[0]);
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
      FILE *f = std::fopen(argv[2], "wb");
      if (!f) {
        std::fprintf(stderr, "cannot write %s\n", argv[2]);
        rc = 1;
      } else {
        // PFM rows run bottom to top
        std::fprintf(f, "PF\n%d %d\n-1.0\n", film->getWidth(), film->getHeight());
        for (int y = film->getHeight() - 1; y >= 0; --y)
          std::fwrite(&img[3 * (size_t)y * film->getWidth()], sizeof(float), 3 * (size_t)film->getWidth(), f);
        std::fclose(f);
      }
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
