// SPPM (SPPM::render, integrator/sppm.cpp:58-322) on gfx950.
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "kernels.h"
#include "runtime.h"

namespace orion_amd {

struct SppmBuffers {
  DevBuf mem;
};

void destroy_sppm(SppmBuffers *b) { delete b; }

}  // namespace orion_amd
