// Kernel launchers shared by the runtime (host side of the HIP library).
#pragma once
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "wavefront.h"

namespace orion_amd {

constexpr int kTraceBlock = 128;  // threads per traversal block (2 wave64)

void launch_raygen(const DScene &sc, const Wave &w, hipStream_t st);
void launch_trace(const DScene &sc, const Wave &w, bool any, int blocks, hipStream_t st);
void launch_between(const Wave &w, int nxt, hipStream_t st);
void launch_shade(const DScene &sc, const Wave &w, int cur, int maxDepth, hipStream_t st);
void launch_film(const Wave &w, int S, float *film, int *count, hipStream_t st);
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, hipStream_t st);

// SPPM (sppm_kernels.hip)
struct SppmState;
int sppm_render(const DScene &sc, SppmState *st, int n_iters, int photons, float r0, int maxDepth,
                int hash_size, int debug_iter, int *vp_cell_count, int *photon_cell_count, int *M,
                int *grid_info, float *grid_bounds, float *film, int *count, unsigned long long *stats,
                hipStream_t stream, int W, int H);

}  // namespace orion_amd
