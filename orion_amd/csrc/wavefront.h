// Wavefront path-tracing state in HBM (structure of arrays, one entry per
// path of the current wave) and the wave64 queue helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orion_amd {

// flags
enum : int {
  F_ALIVE = 1,     // has a continuation ray in ray slot 0
  F_SPEC = 2,      // specularBounce
  F_PEND = 4,      // direct-lighting estimate waiting for its rays
  F_SHADOW = 8,    // ... with a shadow ray in slot 2
  F_MIS = 16,      // ... with a BSDF-sampled (MIS) ray in slot 1
};

// control block (ints, device memory)
enum : int {
  C_PATH0 = 0, C_PATH1 = 1,   // path queue lengths (double buffered)
  C_RAY = 2,                  // closest-hit queue length
  C_SHADOW = 3,               // any-hit queue length
  C_WORK_RAY = 4, C_WORK_SHADOW = 5,
  C_INTS = 16,
};
// 64-bit statistics
enum : int {
  S_CLOSEST = 0, S_SHADOW = 1, S_REFEQ = 2, S_PATHS = 3, S_NODES_C = 4, S_TRIS_C = 5,
  S_PHOTONS = 6, S_DEPOSITS = 7, S_NODES_S = 8, S_TRIS_S = 9, S_COUNT = 12,
};

struct Wave {
  int N;          // paths in the wave
  int npix;       // pixels per sample slab
  uint32_t k0;    // first sample index of the wave
  int W, H;
  // rays: slot s of path p at s * N + p; slots 0 continuation, 1 MIS, 2 shadow
  float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
  // hits for slots 0 and 1
  float *ht, *hu, *hv;
  int *hprim;
  int *vis;       // shadow visibility per path
  // path state
  float *br, *bg, *bb;   // beta
  float *Lr, *Lg, *Lb;   // radiance
  int *dim, *bounce, *flags;
  // pending direct lighting
  float *Ar, *Ag, *Ab;   // light-sampled term (counts when unoccluded)
  float *fr, *fg, *fb;   // BSDF-sampled f * |cos| for the MIS ray
  float *px, *py, *pz;   // shading point
  float *bpdf, *sel;     // bsdf pdf of the MIS direction, light selection pdf
  float *nbr, *nbg, *nbb;  // beta at the shading point
  int *nlight;
  // queues
  int *qpath[2];
  int *qray;      // entries s * N + p, s in {0, 1}
  int *qshadow;   // path ids
  int *ctrl;
  unsigned long long *stats;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Aggregated append: one atomic per wave, slots in lane order.  Must be
// called by every lane of the wave (outside divergent code).
__device__ __forceinline__ int wave_append(int *counter, bool pred) {
  unsigned long long mask = __ballot(pred);
  int lane = lane_id();
  int n = __popcll(mask);
  int base = 0;
  if (n) {
    int leader = __ffsll((long long)mask) - 1;
    if (lane == leader) base = atomicAdd(counter, n);
    base = __shfl(base, leader);
  }
  unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
  return base + __popcll(below);
}

}  // namespace orion_amd
