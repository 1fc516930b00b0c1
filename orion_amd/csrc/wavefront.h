// Wavefront path-tracing state in HBM (one 16-byte record per path in each
// of a few arrays) and the wave64 queue helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_scene.h"

namespace orion_amd {

// flags
enum : int {
  F_ALIVE = 1,     // has a continuation ray in ray slot 0
  F_SPEC = 2,      // specularBounce
  F_PEND = 4,      // direct-lighting estimate waiting for its rays
  F_SHADOW = 8,    // ... with a shadow ray in slot 2
  F_MIS = 16,      // ... with the visibility ray of its BSDF-sampled (MIS) term in slot 1
};

// control block (ints, device memory)
enum : int {
  C_PATH0 = 0, C_PATH1 = 1,   // path queue lengths (double buffered)
  C_RAY = 2,                  // closest-hit queue length
  C_SHADOW = 3,               // any-hit queue length
  C_WORK_RAY = 4, C_WORK_SHADOW = 5,
  C_NPATHS = 6,               // camera paths in the current wave (all jobs)
  C_NPIX = 7,                 // film pixels of all jobs in the wave
  C_NJOBS = 8,                // jobs in the current wave
  C_PQ = 9,                   // (SPPM) photon queue lengths, 3 rotating counters
  C_INTS = 16,
};
// 64-bit statistics, each striped over kStatStripes counters (a block adds
// to stripe blockIdx % kStatStripes) so no single address takes every
// block's atomics; readers sum the stripes.
enum : int {
  S_CLOSEST = 0, S_SHADOW = 1, S_REFEQ = 2, S_PATHS = 3, S_NODES_C = 4, S_TRIS_C = 5,
  S_PHOTONS = 6, S_DEPOSITS = 7, S_NODES_S = 8, S_TRIS_S = 9, S_OVERFLOW = 10, S_SHADED = 11,
  // SPPM photon pass: traversal work, rays, grid lookups, visible points
  // scanned, hits recorded for deposits
  S_NODES_P = 12, S_TRIS_P = 13, S_PHOTON_RAYS = 14, S_LOOKUPS = 15, S_SCANNED = 16, S_PHOTON_HITS = 17,
  S_COUNT = 18,
};
constexpr int kStatStripes = 64;
constexpr size_t kStatBytes = (size_t)S_COUNT * kStatStripes * 8;
constexpr int kSegMaxBounces = 64;  // bounces whose longest segment is recorded (maxDepth is at most 60)
constexpr size_t kSegMaxBytes = (size_t)kSegMaxBounces * 2 * kStatStripes * 8;

// Queue segments: path ids [s * seg, (s + 1) * seg) form segment s.  Each
// segment has its own path, closest-ray (seg entries) and any-hit (2 seg
// entries: shadow and MIS visibility rays) queues with their own lengths, filled by the blocks that shade the
// segment (one device atomic per queue per 256 paths, on the segment's own
// counters: appends never contend on one shared counter).  seg is kSegMin or kSegMax, chosen per wave: waves of 2^21
// paths or more take long segments (a traversal wave's rays then come from
// neighbouring pixels, in one XCD's L2: 4M paths per wave ran 25% faster
// at 16384 than at 512-1024), smaller waves short ones (at 2^18 paths 16
// long segments ran 25% slower than 512 short ones).  Shading balances by
// giving a segment several blocks (Wave::shade_parts).
constexpr int kSegMin = 512, kSegMax = 16384;
inline void choose_segments(int npaths, int *seg, int *nseg) {
  const int s = npaths < (1 << 21) ? kSegMin : kSegMax;
  *seg = s;
  *nseg = npaths > 0 ? (npaths + s - 1) / s : 1;
}

// One render job of a wave (device job table).  A wave packs the next
// sample slabs of any number of jobs -- each with its own scene, camera,
// film, sample range and depth -- into one set of queues; path ids
// [path0, path0 + S * npix) belong to the job, sample s of pixel q at
// path0 + s * npix + q.  In multi-job waves path0 is a multiple of the
// wave's segment size (the path ids up to the next job are padding).
struct Job {
  DScene sc;          // geometry, materials, lights; sc.cam is this job's camera
  int W, H, npix;
  int S;              // samples of this job in the wave
  uint32_t k0;        // sample index of the job's first sample in the wave
  int path0;          // first path id of the job
  int pix0;           // first film pixel of the job (prefix over jobs)
  int maxDepth;
  float *film;        // W * H * 3 sums
  int *count;         // W * H sample counts
};

struct Wave {
  int N;          // slot stride: capacity of the per-path arrays
  int npix;       // (SPPM) pixels of the image
  uint32_t k0;
  int W, H;
  // Per-path state as float4 records, so that a path's fields are a few
  // 16-byte loads and stores, each in one cache line: shading and traversal
  // reach paths through compacted queues, and a scalar array per field put
  // every field of a path in a line of its own (shading is bound by the
  // lines its L1 can have in flight).
  //   rays, slot s of path p at s * N + p (slots: 0 continuation, 1 MIS
  //   visibility, 2 shadow): ro = (origin, tmax), rd = (direction, tie as
  //   int bits: the MIS visibility query's tie primitive, kTieAll otherwise)
  float4 *ro, *rd;
  float4 *hit;    // closest hit of slot 0: (t, u, v, primitive as int bits; -1 = miss)
  int *vis;       // any-hit results: vis[p] shadow ray, vis[N + p] MIS ray (1 = unoccluded)
  float4 *st;     // (beta, path_bits(flags, bounce, Sobol dimension))
  float4 *rad;    // (L, unused)
  // pending direct lighting: (A, light selection pdf), (M, unused), (beta at
  // the shading point, unused); A counts when the shadow ray is unoccluded,
  // M (the BSDF-sampled MIS term) when its visibility ray is
  float4 *pa, *pm, *pn;
  // queues
  int *qpath[2];
  int *qray;      // path ids (slot 0 rays)
  int *qshadow;   // p: shadow ray (slot 2); N + p: MIS visibility ray (slot 1)
  int *ctrl;
  unsigned long long *stats;
  // Longest queue segment of each bounce (kSegMaxBounces x {closest-hit,
  // any-hit}, striped like the statistics), recorded by the traversal when
  // non-null: the wave runner sizes later waves' traversal grids from it.
  unsigned long long *segmax;
  // traversal stack spill area: spill_cap entries per traversal thread
  int *spill;
  int spill_cap;
  // queue segments: `seg` paths per segment (any-hit queues hold 2 seg
  // entries per segment), nseg segments in use, lengths per segment
  int seg, nseg;
  int *cnt_ray, *cnt_shadow;  // the lengths a traversal launch reads
  // Path pipeline queue lengths, three rotating sets: bounce b traces the
  // ray lengths of set b % 3, shades the path lengths of set b % 3, appends
  // to set (b + 1) % 3 and zeroes set (b + 2) % 3 (nothing reads it during
  // bounce b).  Set t: path lengths at cnt + 3 t cnt_stride, closest-ray
  // lengths cnt_stride after them, shadow lengths 2 cnt_stride after.
  int *cnt;
  int cnt_stride;
  int shade_parts;  // shading blocks per segment

  const int *seg_job;  // multi-job waves: job index of each segment
  // multi-job waves: the segment behind each logical segment index.  Blocks
  // map to logical indices XCD-contiguously (xcd_remap); the order deals
  // every job's segments to the eight XCDs in eight contiguous parts, so
  // each XCD gets a share of every job (balanced) and neighbouring segments
  // (one L2).  One-job waves use the identity.
  const int *seg_perm;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// st.w: flags (bits 0-7), bounce (8-15), next Sobol dimension (16-31)
__device__ __forceinline__ float path_bits(int flags, int bounce, int dim) {
  return __int_as_float(flags | (bounce << 8) | (dim << 16));
}
__device__ __forceinline__ int bits_flags(float4 s) { return __float_as_int(s.w) & 0xff; }
__device__ __forceinline__ int bits_bounce(float4 s) { return (__float_as_int(s.w) >> 8) & 0xff; }
__device__ __forceinline__ int bits_dim(float4 s) { return (int)((uint32_t)__float_as_int(s.w) >> 16); }
// ray slot `slot` (s * N + p)
__device__ __forceinline__ void put_ray(const Wave &w, size_t slot, dev::V3 o, dev::V3 d, float tmax,
                                        int tie = dev::kTieAll) {
  w.ro[slot] = make_float4(o.x, o.y, o.z, tmax);
  w.rd[slot] = make_float4(d.x, d.y, d.z, __int_as_float(tie));
}

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD
// b % 8) and each XCD has its own L2.  xcd_block maps the hardware block id
// to a logical one so that each XCD takes a contiguous range of logical
// blocks: neighbouring queue segments (neighbouring pixels, coherent rays)
// then share one L2 instead of being fetched into all eight.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int per = G / kXcds, rem = G % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x * per + min(x, rem) + k;
}
__device__ __forceinline__ int xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }

__device__ __forceinline__ void stat_add(unsigned long long *stats, int s, unsigned long long v) {
  if (stats && v) atomicAdd(&stats[s * kStatStripes + (blockIdx.x & (kStatStripes - 1))], v);
}

// Wave-aggregated append of pred lanes to an LDS counter; returns each
// lane's slot (lane order).  Every lane of the wave must call it.
__device__ __forceinline__ int lds_append(int *counter, bool pred) {
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

// Block-aggregated append: one device atomic per block, slots in thread
// order.  Returning atomics on one address serialise at L2 (~10 ns each), so
// a 2^20-entry queue filled one atomic per wave costs ~160 us; per block of
// 256 threads it is a quarter of that.  Every thread of the block must call
// it (uniform control flow); lds holds blockDim.x / 64 + 1 ints.
__device__ __forceinline__ int block_append(int *counter, bool pred, int *lds) {
  unsigned long long mask = __ballot(pred);
  const int lane = lane_id(), wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) lds[wv] = __popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int v = 0; v < nw; ++v) { int c = lds[v]; lds[v] = tot; tot += c; }
    lds[nw] = tot ? atomicAdd(counter, tot) : 0;
  }
  __syncthreads();
  unsigned long long below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
  int r = lds[nw] + lds[wv] + __popcll(below);
  __syncthreads();  // lds is free for the next call
  return r;
}

// Block-aggregated appends to four queues at once: one device atomic per
// queue and block (counters may repeat: each append gets its own range),
// three barriers.  slot[k] is the thread's slot in queue k where pk holds.
// lds: 4 * (waves per block + 1) ints; every thread of the block must call it.
__device__ __forceinline__ void block_append4(int *c0, int *c1, int *c2, int *c3, bool p0, bool p1, bool p2, bool p3,
                                              int *lds, int slot[4]) {
  const int lane = lane_id(), wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned long long m[4] = {__ballot(p0), __ballot(p1), __ballot(p2), __ballot(p3)};
  if (lane == 0)
    for (int k = 0; k < 4; ++k) lds[k * (nw + 1) + wv] = __popcll(m[k]);
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    int *const c = k == 0 ? c0 : k == 1 ? c1 : k == 2 ? c2 : c3;
    int *const l = lds + k * (nw + 1);
    int tot = 0;
    for (int v = 0; v < nw; ++v) { const int n = l[v]; l[v] = tot; tot += n; }
    l[nw] = tot ? atomicAdd(c, tot) : 0;
  }
  __syncthreads();
  const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (int k = 0; k < 4; ++k) slot[k] = lds[k * (nw + 1) + nw] + lds[k * (nw + 1) + wv] + __popcll(m[k] & below);
  __syncthreads();  // lds is free for the next call
}

}  // namespace orion_amd
