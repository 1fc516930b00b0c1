// Kernel launchers shared by the runtime (host side of the HIP library).
#pragma once
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "wavefront.h"

namespace orion_amd {

constexpr int kTraceBlock = 128;  // threads per traversal block (2 wave64)

// Wave launchers: `jobs` is the device job table, `multi` selects the
// per-path job lookup (false: every path belongs to job 0).
void launch_raygen(const Job *jobs, const Wave &w, int max_paths, bool multi, hipStream_t st);
void launch_trace(const Job *jobs, const Wave &w, bool any, int stack_entries, int blocks, bool multi,
                  hipStream_t st);
int stack_variant(int bvh_depth);
int trace_blocks_per_cu(int stack_entries, bool any);
void launch_between(const Wave &w, int nxt, hipStream_t st);
void launch_shade(const Job *jobs, const Wave &w, int cur, int max_paths, bool multi, hipStream_t st);
void launch_film(const Job *jobs, const Wave &w, int max_pix, bool multi, hipStream_t st);
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, hipStream_t st);

}  // namespace orion_amd
