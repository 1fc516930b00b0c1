// Runtime objects behind the C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>

#include <string>

#include "../../include/orion_amd.h"
#include "bvh.h"
#include "device_scene.h"
#include "wavefront.h"

namespace orion_amd {

int dev_alloc(void **p, size_t bytes);

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  int ensure(size_t b);
  void release();
};

struct WaveBuffers {
  DevBuf mem;
  int cap = 0;
  Wave wave{};
  int ensure(int N);
};

struct SppmBuffers;
void destroy_sppm(SppmBuffers *b);
struct WaveRunner;
void destroy_runner(WaveRunner *r);
// drops the runner's traversal-grid plan when it was made for `scene`
void forget_scene(WaveRunner *r, const void *scene);
// sums the striped device statistics (kStatBytes at dev) into out[S_COUNT]
int read_striped_stats(const void *dev, unsigned long long *out);
// ORION_ERR_UNSUPPORTED when a traversal dropped stack entries (a tree
// deeper than its stack bound: the hits of that call are not trustworthy)
int check_overflow(unsigned long long dropped);
// Largest wave (paths in flight): ray slot s of path p lives at s * N + p
// with s < 3, so 3 N must stay below 2^31.
constexpr int kMaxWavePathsLimit = 1 << 28;

}  // namespace orion_amd

struct orion_ctx_s {
  int device = 0;
  int n_cu = 0;
  std::string arch;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  orion_amd::WaveRunner *runner = nullptr;  // wave buffers of every render on this stream
  // this device's copy of the allocation-free intersection's scratch array
  // (its device address, from hipGetSymbolAddress at context creation)
  int *scratch = nullptr;
  // scenes and servers created on this context and not yet destroyed: the
  // context refuses to go before them (they use its stream and runner)
  std::atomic<int> children{0};
};

struct orion_scene_s {
  orion_ctx ctx = nullptr;
  int W = 0, H = 0;
  orion_amd::Bvh bvh;
  orion_amd::DevBuf nodes, tris, shade, mats, lights, light_cdf, light_func, light_area, sobol;
  orion_amd::DevBuf env_tex, env_dist, env_meta;
  orion_amd::DevBuf film_avg;  // getAveraged output
  orion_amd::DevBuf film, count, ctrl, stats;
  orion_amd::DevBuf spill;  // traversal stack spill area of the intersection API
  orion_amd::DScene d{};
  orion_amd::WaveBuffers wave;
  orion_amd::SppmBuffers *sppm = nullptr;
  double build_ms = 0, upload_ms = 0;
  ~orion_scene_s();
};
