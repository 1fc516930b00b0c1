// Kernel launchers shared by the runtime (host side of the HIP library).
#pragma once
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "wavefront.h"

namespace orion_amd {

constexpr int kTraceBlock = dev::kRingStride;  // threads per traversal block (one wave64)
constexpr int kDefaultRing = 16;          // LDS ring entries per lane (8, 16 or 32)
// Queue entries per traversal wave64.  One-job waves of at least
// kLargeWavePaths paths take kLargeWaveChunk (a 2^26-path wave of the 1.3M-
// triangle scene: 256 and 1024 measured 2% and 4% slower, 128 9%); smaller
// and multi-job waves kDefaultChunk (jobs of 65k-67M paths on 1k-500k
// triangles rendered one at a time: 512 for all of them 9% slower).
constexpr int kDefaultChunk = 256;
constexpr int kLargeWaveChunk = 512;
constexpr int kLargeWavePaths = 1 << 25;
constexpr int kShadeThreads = 256;        // threads per path-shading block

// Wave launchers: `jobs` is the device job table, `multi` selects the
// per-path job lookup (false: every path belongs to job 0).
void launch_raygen(const Job *jobs, const Wave &w, bool multi, hipStream_t st);
// ring: LDS ring entries per lane (8, 16 or 32); chunk: queue entries per
// wave64.  w.spill must hold w.spill_cap entries for each of the launch's
// w.nseg * cps * kTraceBlock threads (at most trace_grid(w.nseg, w.seg,
// any, chunk) blocks).  cps: chunk positions per segment, each wave taking
// every cps-th chunk of its segment (<= 0: all of a full segment's chunks,
// one wave per chunk).  rec >= 0: record the bounce's longest segment in
// w.segmax slot rec.
void launch_trace(const Job *jobs, const Wave &w, bool any, int ring, int chunk, bool multi, hipStream_t st,
                  int cps = 0, int rec = -1);
// an SPPM photon bounce (closest hit, trace_grid(w.nseg, w.seg, false,
// chunk) blocks, or w.nseg * cps with cps > 0; counted apart from camera and
// path rays); rec >= 0: record the queue's length in w.segmax slot rec
void launch_trace_photons(const Job *jobs, const Wave &w, int ring, int chunk, hipStream_t st, int cps = 0,
                          int rec = -1);
// closest-hit and any-hit queues in one launch (trace_pair_grid blocks, or
// fewer with cps_c / cps_a > 0 chunk positions per segment for the two
// queues); rec >= 0: record the two queues' lengths in w.segmax slots 2 rec
// and 2 rec + 1
void launch_trace_pair(const Job *jobs, const Wave &w, int ring, int chunk, bool multi, hipStream_t st,
                       int cps_c = 0, int cps_a = 0, int rec = -1);
int trace_pair_grid(int nseg, int seg, int chunk);
int ring_variant(int ring);
int trace_chunk(int chunk);
// blocks of kTraceBlock threads covering nseg queue segments (closest-hit
// queues: seg entries per segment; any-hit: 2 seg)
int trace_grid(int nseg, int seg, bool any, int chunk);
// spill area (ints) for `blocks` traversal blocks; the per-thread capacity
// is max(0, stack_bound - ring)
size_t trace_spill_ints(int stack_bound, int ring, int blocks);
int trace_waves_per_simd(int ring);  // occupancy of the closest-hit traversal kernel
// single-queue waves (SPPM): account the traced rays, reset the queues
void launch_between(const Wave &w, int nxt, hipStream_t st);
void launch_shade(const Job *jobs, const Wave &w, int cur, bool multi, hipStream_t st);
void launch_film(const Job *jobs, const Wave &w, int max_pix, bool multi, hipStream_t st);
// closest hits on `blocks` one-wave blocks striding over the rays' chunks of
// kDefaultChunk; spill holds spill_cap entries for each of blocks x
// kTraceBlock threads (with the default LDS ring); *overflow counts dropped
// stack entries
void launch_intersect_fixed(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *spill, int spill_cap, int blocks, unsigned int *overflow,
                            hipStream_t st);
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, int *spill, int spill_cap,
                            unsigned long long *stats, hipStream_t st);

}  // namespace orion_amd
