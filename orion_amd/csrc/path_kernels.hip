// Wavefront path tracer for gfx950 (PathTracing::Li, integrator/path.cpp:50-99,
// with direct lighting from core/integrator.cpp:8-90).
//
// Per wave: raygen -> [closest-hit trace (continuation + MIS rays) ->
// any-hit trace (shadow rays) -> shade] x (maxDepth + 1) -> film.  The shade
// kernel first resolves the direct-lighting estimate pending from the
// previous bounce (so L accumulates in the reference's order: emission of
// bounce b, direct light of bounce b, emission of bounce b + 1, ...), then
// shades the new hit: emission, light sampling, BSDF sampling, Russian
// roulette.  Queues are compacted with wave64 ballots, one atomic per block
// step on the queue segment's own counters.
//
// A wave may pack several jobs (device job table, wavefront.h): every path
// carries its job index and reads that job's scene, camera, depth and film.
// MULTI = false is the one-job fast path (job 0, no per-path job reads).
// Per-wave sizes live in the control block, so a wave's launch sequence has
// fixed arguments and can be captured once as a hipGraph.
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "kernels.h"
#include "shading.h"
#include "wavefront.h"

namespace orion_amd {
using namespace dev;

// One thread per path slot of the wave's segments: camera rays for the
// wave's paths, and every segment's initial queue lengths (its live paths in
// the path and closest queues, none in the shadow queue).  Multi-job waves
// start every job at a segment boundary, so a segment belongs to one job and
// its live paths are a prefix of it (the job's padding follows).
template <bool MULTI>
__global__ void k_raygen(const Job *__restrict__ jobs, Wave w) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= w.nseg * w.seg) return;
  const Job &J = jobs[MULTI ? w.seg_job[p / w.seg] : 0];
  const int live_end = J.path0 + J.S * J.npix;
  if (p % w.seg == 0) {
    // set 0: this bounce's lengths; set 1: empty, bounce 0 appends to it
    const int sg = p / w.seg, cs = w.cnt_stride;
    const int n = max(0, min(w.seg, live_end - p));
    w.cnt[sg] = n;
    w.cnt[cs + sg] = n;
    w.cnt[2 * cs + sg] = 0;
    w.cnt[3 * cs + sg] = 0;
    w.cnt[4 * cs + sg] = 0;
    w.cnt[5 * cs + sg] = 0;
    stat_add(w.stats, S_CLOSEST, (unsigned long long)n);
  }
  if (p >= live_end) return;
  int q = p - J.path0;
  int s = q / J.npix, pix = q - s * J.npix;
  int j = pix / J.W, i = pix - j * J.W;
  uint32_t index = (J.k0 + (uint32_t)s) * (uint32_t)J.npix + (uint32_t)pix;  // path.cpp:26
  V3 o, d;
  camera_ray(J.sc, i, j, index, &o, &d);
  put_ray(w, p, o, d, kFltMax);  // bounce 0's shading starts the path state
  const int sg = p / w.seg, j0 = p - sg * w.seg;
  w.qpath[0][p] = p;
  w.qray[sg * w.seg + j0] = p;
}

// ---------------------------------------------------------------- traversal
// Wave queues as traversal-engine IO.  Closest rays: path ids, ray in
// slot 0.  Any-hit entries r < N: shadow ray of path r (slot 2); N + p: MIS
// visibility ray of path p (slot 1, its tie primitive in hprim[N + p]).
// The engine's entry q is the segment's entry (c + (q / chunk) cps) chunk +
// q mod chunk: the wave takes the segment's chunks c, c + cps, c + 2 cps ...
template <bool ANY>
struct WaveRays {
  const Wave *w;
  int base;                   // first queue entry of the segment
  int c, cps, chunk;          // the wave's first chunk, chunk stride, chunk size
  const float4 *nodes, *tris;  // the segment's job's scene (wave-uniform)
  __device__ int fetch(int q) const {
    const int k = q / chunk;
    const int e = (c + k * cps) * chunk + (q - k * chunk);
    return (ANY ? w->qshadow : w->qray)[base + e];
  }
  __device__ void load(int r, TRay *ray) const {
    const int slot = ANY && r < w->N ? 2 * w->N + r : r;
    const float4 a = w->ro[slot], b = w->rd[slot];
    ray->o = mk(a.x, a.y, a.z);
    ray->d = mk(b.x, b.y, b.z);
    ray->tmin = 0.0f;
    ray->tmax = a.w;
    ray->tag = r;
    ray->tie = ANY ? __float_as_int(b.w) : kTieAll;
  }
  __device__ void store(const TRay &ray, bool hit, const HitRec &h) const {
    const int r = ray.tag;
    if (ANY) w->vis[r] = hit ? 0 : 1;
    else w->hit[r] = hit ? make_float4(h.t, h.u, h.v, __int_as_float(h.prim)) : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
  }
};

// A launch part covers `cps` chunk positions of every queue segment: wave64
// number g (its block's logical index within the part, XCD-contiguous)
// takes segment g / cps, chunks c, c + cps, c + 2 cps ... (c = g mod cps) of
// `chunk` entries, up to the segment's length.  cps = the chunks of a full
// segment gives each wave at most one chunk; a smaller one (a wave runner
// that knows a bounce's segments are short) saves dispatching waves that
// would find nothing, and still traces every entry.  rec >= 0: the wave of
// chunk position 0 records the segment's length as the bounce's longest
// (w.segmax slot rec).  `ring` is the block's LDS.  PHOTON: an SPPM photon
// bounce (closest hit; its traversal work has its own counters).
template <bool ANY, int RING, bool MULTI, bool PHOTON = false>
__device__ __forceinline__ void trace_block(const Job *jobs, const Wave &w, int chunk, int cps, int rec, int lb,
                                            int *ring) {
  static_assert(kTraceBlock == 64, "one wave64 per traversal block");
  const int seg_entries = ANY ? 2 * w.seg : w.seg;  // the queue's stride per segment
  if (cps <= 0) cps = (seg_entries + chunk - 1) / chunk;
  // the wave's index is uniform (one wave per block), so everything derived
  // from it (segment, job, scene pointers) stays in scalar registers
  const int ls = lb / cps;
  if (ls >= w.nseg) return;
  const int c = lb - ls * cps;
  const int sg = MULTI ? w.seg_perm[ls] : ls;
  const int count = (ANY ? w.cnt_shadow : w.cnt_ray)[sg];
  if (rec >= 0 && c == 0 && count > 0 && w.segmax && __lane_id() == 0)
    atomicMax(&w.segmax[(size_t)rec * kStatStripes + (blockIdx.x & (kStatStripes - 1))], (unsigned long long)count);
  const int nch = (count + chunk - 1) / chunk;  // the segment's chunks
  if (c >= nch) return;
  const int m = (nch - 1 - c) / cps + 1;        // the wave's chunks
  const int last = c + (m - 1) * cps;
  const int n = (m - 1) * chunk + (last == nch - 1 ? count - last * chunk : chunk);
  TravStack<RING> stk;
  stk.ring = ring + threadIdx.x;
  stk.spill_stride = gridDim.x * kTraceBlock;
  stk.spill = w.spill + (size_t)blockIdx.x * kTraceBlock;
  stk.spill_cap = w.spill_cap;
  uint32_t ovf = 0, nv = 0, nt = 0;
  const DScene &sc = jobs[MULTI ? w.seg_job[sg] : 0].sc;
  WaveRays<ANY> io{&w, sg * seg_entries, c, cps, chunk, sc.nodes, sc.tris};
  trace_engine<ANY, RING>(io, 0, n, stk, &nv, &nt, &ovf);
  const int lane = (int)__lane_id();
  // nv / nt are the wave's totals (for the roofline's n_node / n_tri)
  if (ovf) stat_add(w.stats, S_OVERFLOW, ovf);
  if (lane == 0) {
    stat_add(w.stats, ANY ? S_NODES_S : PHOTON ? S_NODES_P : S_NODES_C, nv);
    stat_add(w.stats, ANY ? S_TRIS_S : PHOTON ? S_TRIS_P : S_TRIS_C, nt);
  }
}

template <bool ANY, int RING, bool MULTI>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_trace(const Job *jobs, Wave w, int chunk, int cps, int rec) {
  __shared__ int ring[trace_lds_ints<RING>()];
  trace_block<ANY, RING, MULTI>(jobs, w, chunk, cps, rec, xcd_remap(blockIdx.x, gridDim.x), ring);
}

// An SPPM photon bounce's closest-hit traversal: the same engine, its own
// counters and kernel (profiles attribute it apart from the path tracer's).
template <int RING>
__global__ __launch_bounds__(kTraceBlock) void k_trace_photons(const Job *jobs, Wave w, int chunk, int cps, int rec) {
  __shared__ int ring[trace_lds_ints<RING>()];
  trace_block<false, RING, false, true>(jobs, w, chunk, cps, rec, xcd_remap(blockIdx.x, gridDim.x), ring);
}

// Closest-hit and any-hit queues of a bounce in one launch: blocks
// [0, n_closest) take the closest-hit part, the rest the any-hit part, so
// the any-hit blocks fill the GPU while the closest-hit part drains (and
// the other way round).  n_closest is a multiple of the XCD count, so a
// block's XCD is the same in its part's numbering.
template <int RING, bool MULTI>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_trace_pair(
    const Job *jobs, Wave w, int chunk, int n_closest, int cps_c, int cps_a, int rec) {
  __shared__ int ring[trace_lds_ints<RING>()];
  const int b = blockIdx.x;
  if (b < n_closest)
    trace_block<false, RING, MULTI>(jobs, w, chunk, cps_c, rec >= 0 ? 2 * rec : -1, xcd_remap(b, n_closest), ring);
  else
    trace_block<true, RING, MULTI>(jobs, w, chunk, cps_a, rec >= 0 ? 2 * rec + 1 : -1,
                                   xcd_remap(b - n_closest, gridDim.x - n_closest), ring);
}

__global__ void k_between(Wave w, int nxt) {
  // runs after both traces, before shade: account the rays just traced and
  // reset the queues the shade kernel appends to
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    stat_add(w.stats, S_CLOSEST, (unsigned long long)w.ctrl[C_RAY]);
    stat_add(w.stats, S_SHADOW, (unsigned long long)w.ctrl[C_SHADOW]);
    w.ctrl[C_RAY] = 0;
    w.ctrl[C_SHADOW] = 0;
    w.ctrl[C_WORK_RAY] = 0;
    w.ctrl[C_WORK_SHADOW] = 0;
    w.ctrl[C_PATH0 + nxt] = 0;
  }
}

// ---------------------------------------------------------------- shade
// Bounce b: the w.shade_parts blocks of segment sg shade consecutive parts
// of its live paths (set b % 3 lengths) and append to the segment's next
// path, closest-ray and shadow-ray queues (set (b + 1) % 3 lengths) with one
// device atomic per queue per 256 paths.  Several blocks per segment keep
// segments long -- the traversal's rays come from neighbouring pixels --
// while the shading launch still has enough blocks to balance the GPU.
// `jobs` is __restrict__ (nothing writes the job table during a launch): the
// scene's pointers are then read through the scalar cache, where otherwise
// every store to the wave's arrays forced a vector reload of each pointer
// before its next dependent load (measured 8% faster shading).
template <bool MULTI>
__global__ __launch_bounds__(kShadeThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_shade(const Job *__restrict__ jobs, Wave w, int it) {
  __shared__ int lw[kShadeThreads / 64][3];  // per wave64: its appends, then their offsets in the block's range
  __shared__ int lbase[3];  // the block's ranges in the three queues
  const int cur = it & 1, nxt = cur ^ 1;
  const int K = w.shade_parts, cs = w.cnt_stride;
  const int lb = xcd_block();
  const int ls = lb / K, part = lb - ls * K;
  const int sg = MULTI ? w.seg_perm[ls] : ls;
  const int seg0 = sg * w.seg;
  int *const cnt_cur = w.cnt + 3 * cs * (it % 3);
  int *const cnt_nxt = w.cnt + 3 * cs * ((it + 1) % 3);
  if (part == 0 && threadIdx.x < 3) w.cnt[3 * cs * ((it + 2) % 3) + threadIdx.x * cs + sg] = 0;
  const int count = cnt_cur[sg];
  const int per = ((count + K - 1) / K + kShadeThreads - 1) / kShadeThreads * kShadeThreads;
  const int p_begin = min(count, part * per), p_end = min(count, p_begin + per);
  const int N = w.N;
  const Job &J = jobs[MULTI ? w.seg_job[sg] : 0];  // the segment's job
  const DScene &sc = J.sc;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t n_done = 0, n_mis = 0, n_ray = 0, n_shadow = 0;
  for (int base = p_begin; base < p_end; base += blockDim.x) {
    int i = base + threadIdx.x;
    bool act = i < p_end;
    int p = act ? w.qpath[cur][seg0 + i] : 0;
    bool want_ray0 = false, want_ray1 = false, want_shadow = false, keep = false, done = false;
    int ref_rays = 0;  // closest-hit casts orion makes for the MIS sample
    if (act) {
      // bounce 0 shades the camera paths: their state is the initial one
      // (raygen does not write it)
      const float4 S = it == 0 ? make_float4(1.f, 1.f, 1.f, path_bits(F_ALIVE, 0, 7)) : w.st[p];
      int fl = bits_flags(S), b = bits_bounce(S), dim = bits_dim(S);
      Sp beta = sp3(S.x, S.y, S.z);
      const float4 R = it == 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : w.rad[p];
      Sp L = sp3(R.x, R.y, R.z);
      // 1) resolve the direct-lighting estimate of the previous bounce
      if (fl & F_PEND) {
        resolve_direct(w, p, fl, &L);
        fl &= ~(F_PEND | F_SHADOW | F_MIS);
      }
      // 2) the hit of this bounce
      if (fl & F_ALIVE) {
        const float4 H = w.hit[p];
        const int hp = __float_as_int(H.w);
        const float4 D = w.rd[p];
        V3 d = mk(D.x, D.y, D.z);
        Isect is;
        if (hp >= 0) make_isect(sc, hp, H.y, H.z, &is);
        if ((b == 0 || (fl & F_SPEC)) && hp >= 0 && is.light >= 0) acc(L, beta * area_emission(sc, is, neg(d)));
        if ((b == 0 || (fl & F_SPEC)) && hp < 0) add_env_le(sc, d, beta, &L);  // path.cpp:65-68
        if (hp < 0 || b >= J.maxDepth) {
          fl &= ~F_ALIVE;
        } else {
          Bsdf bs;
          make_bsdf(sc.mats + ORION_MAT_FLOATS * (size_t)is.mat, is.ng, is.ns, is.u, is.v, false, &bs);
          int q = p - J.path0;
          int s = q / J.npix;
          uint32_t index = (J.k0 + (uint32_t)s) * (uint32_t)J.npix + (uint32_t)(q - s * J.npix);
          if (bsdf_num(bs, ALL & ~SPEC) > 0) {
            int nf = setup_direct(sc, w, p, bs, is, d, beta, index, &dim, false, &ref_rays);
            fl |= nf;
            want_shadow = (nf & F_SHADOW) != 0;
            want_ray1 = (nf & F_MIS) != 0;
          }
          // BSDF sampling for the continuation (path.cpp:81-96)
          float u1 = sobol(sc.sobol, index, dim++), u0 = sobol(sc.sobol, index, dim++);
          V3 wi = mk(0, 0, 0);
          float pdf = 0;
          int type = 0;
          Sp f = bsdf_sample(bs, &wi, neg(d), u0, u1, &pdf, ALL, &type);
          if (black(f) || pdf == 0) {
            fl &= ~F_ALIVE;
          } else {
            mulby(beta, f * absdot(wi, is.ns) / pdf);
            fl = (type & SPEC) ? (fl | F_SPEC) : (fl & ~F_SPEC);
            V3 o = offset_ray_origin(is.p, is.ng, wi);
            bool alive = true;
            if (b > 3) {
              float qq = mx(0.05f, 1 - maxc(beta));
              if (sobol(sc.sobol, index, dim++) < qq) alive = false;
              else beta = beta / (1 - qq);
            }
            if (alive) {
              put_ray(w, p, o, wi, kFltMax);
              ++b;
              want_ray0 = true;
            } else {
              fl &= ~F_ALIVE;
            }
          }
        }
      }
      w.rad[p] = make_float4(L.r, L.g, L.b, 0.f);
      w.st[p] = make_float4(beta.r, beta.g, beta.b, path_bits(fl, b, dim));
      keep = (fl & (F_ALIVE | F_PEND)) != 0;
      done = !keep;
    }
    // appends: per wave64 ballots, per block one atomic per queue
    const uint64_t mk = __ballot(keep), m0 = __ballot(want_ray0), m1 = __ballot(want_ray1),
                   ms = __ballot(want_shadow);
    if (lane == 0) {
      lw[wv][0] = __popcll(mk);
      lw[wv][1] = __popcll(m0);
      lw[wv][2] = __popcll(ms) + __popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int q = threadIdx.x;
      int tot = 0;
      for (int v = 0; v < (int)(blockDim.x >> 6); ++v) {
        const int c = lw[v][q];
        lw[v][q] = tot;
        tot += c;
      }
      lbase[q] = tot ? atomicAdd(&cnt_nxt[q * cs + sg], tot) : 0;
      if (q == 1) n_ray += tot;
      if (q == 2) n_shadow += tot;
    }
    __syncthreads();
    if (keep) w.qpath[nxt][seg0 + lbase[0] + lw[wv][0] + __popcll(mk & below)] = p;
    if (want_ray0) w.qray[seg0 + lbase[1] + lw[wv][1] + __popcll(m0 & below)] = p;
    const int a0 = 2 * seg0 + lbase[2] + lw[wv][2];
    if (want_shadow) w.qshadow[a0 + __popcll(ms & below)] = p;
    if (want_ray1) w.qshadow[a0 + __popcll(ms) + __popcll(m1 & below)] = N + p;
    n_done += done ? 1 : 0;
    // orion casts the MIS sample's ray and counts it twice; the visibility
    // rays traced here are counted as any-hit rays
    n_mis += 2 * ref_rays - (want_ray1 ? 1 : 0);
    __syncthreads();  // lw / lbase are rewritten by the next step
  }
  // block totals: queue lengths of the segment and striped statistics
  for (int off = 32; off > 0; off >>= 1) {
    n_done += __shfl_down(n_done, off);
    n_mis += __shfl_down(n_mis, off);
  }
  __shared__ uint32_t ltot[2];
  if (threadIdx.x < 2) ltot[threadIdx.x] = 0;
  __syncthreads();
  if (lane_id() == 0) {
    atomicAdd(&ltot[0], n_done);
    atomicAdd(&ltot[1], n_mis);
  }
  __syncthreads();
  if (threadIdx.x == 1) stat_add(w.stats, S_CLOSEST, (unsigned long long)n_ray);
  if (threadIdx.x == 2) stat_add(w.stats, S_SHADOW, (unsigned long long)n_shadow);
  if (threadIdx.x == 0) {
    stat_add(w.stats, S_PATHS, ltot[0]);
    stat_add(w.stats, S_REFEQ, ltot[1]);
    stat_add(w.stats, S_SHADED, (unsigned long long)(p_end - p_begin));
  }
}

// RenderTarget::acumulate in sample order per pixel (rendertarget.cpp:22-27)
template <bool MULTI>
__global__ void k_film(const Job *__restrict__ jobs, Wave w) {
  const int total = w.ctrl[C_NPIX];
  const int njobs = w.ctrl[C_NJOBS];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    int jb = 0;
    if (MULTI) {
      int lo = 0, hi = njobs - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].pix0 <= g) lo = mid;
        else hi = mid - 1;
      }
      jb = lo;
    }
    const Job &J = jobs[jb];
    int pix = g - J.pix0;
    float r = J.film[3 * pix], gg = J.film[3 * pix + 1], b = J.film[3 * pix + 2];
    for (int s = 0; s < J.S; ++s) {
      int p = J.path0 + s * J.npix + pix;
      const float4 R = w.rad[p];
      r += R.x;
      gg += R.y;
      b += R.z;
    }
    J.film[3 * pix] = r;
    J.film[3 * pix + 1] = gg;
    J.film[3 * pix + 2] = b;
    J.count[pix] += J.S;
  }
}

// ---------------------------------------------------------------- batch intersection API
struct SoaRays {
  const float4 *nodes, *tris;
  int n;
  const float *rays;
  float *t, *u, *v;
  int *prim, *tie, *occluded;
  __device__ int fetch(int i) const { return i; }
  __device__ void load(int i, TRay *ray) const {
    ray->o = mk(rays[i], rays[n + i], rays[2 * n + i]);
    ray->d = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
    ray->tmin = rays[6 * n + i];
    ray->tmax = rays[7 * n + i];
    ray->tag = i;
    ray->tie = kTieAll;
  }
  __device__ void store(const TRay &ray, bool hit, const HitRec &h) const {
    const int i = ray.tag;
    if (occluded) {
      occluded[i] = hit ? 1 : 0;
    } else {
      t[i] = hit ? h.t : 0.f;
      prim[i] = hit ? h.prim : -1;
      if (u) u[i] = hit ? h.u : 0.f;
      if (v) v[i] = hit ? h.v : 0.f;
      if (tie) tie[i] = hit ? h.tie : 0;
    }
  }
};

template <bool ANY, int RING>
__global__ __launch_bounds__(kTraceBlock) void k_intersect_batch(DScene sc, int n, const float *rays, float *t,
                                                                 int *prim, float *u, float *v, int *tie,
                                                                 int *occluded, int *spill, int spill_cap,
                                                                 int chunk, unsigned long long *stats) {
  __shared__ int ring[trace_lds_ints<RING>()];
  const int begin = (blockIdx.x * (kTraceBlock / 64) + (int)(threadIdx.x >> 6)) * chunk;
  if (begin >= n) return;
  TravStack<RING> stk;
  stk.ring = ring + threadIdx.x;
  stk.spill_stride = gridDim.x * kTraceBlock;
  stk.spill = spill + (size_t)blockIdx.x * kTraceBlock;
  stk.spill_cap = spill_cap;
  uint32_t ovf = 0, nv = 0, nt = 0;
  SoaRays io{sc.nodes, sc.tris, n, rays, t, u, v, prim, tie, ANY ? occluded : nullptr};
  trace_engine<ANY, RING>(io, begin, min(begin + chunk, n), stk, &nv, &nt, &ovf);
  if (ovf) stat_add(stats, S_OVERFLOW, ovf);
}

// Closest hits of n rays on a fixed grid: the wave64 of block b traces the
// ray chunks b, b + G, b + 2G, ... (G = gridDim.x), so the spill area is G x
// 64 lanes x spill_cap entries whatever n is.  Dropped stack entries are
// added to *overflow.
static_assert(kTraceBlock == 64, "one wave64 per traversal block");
template <int RING>
__global__ __launch_bounds__(kTraceBlock) void k_intersect_fixed(DScene sc, int n, const float *rays, float *t,
                                                                 int *prim, float *u, float *v, int *tie,
                                                                 int *spill, int spill_cap, int chunk,
                                                                 unsigned int *overflow) {
  __shared__ int ring[trace_lds_ints<RING>()];
  TravStack<RING> stk;
  stk.ring = ring + threadIdx.x;
  stk.spill_stride = gridDim.x * kTraceBlock;
  stk.spill = spill + (size_t)blockIdx.x * kTraceBlock;
  stk.spill_cap = spill_cap;
  uint32_t ovf = 0, nv = 0, nt = 0;
  SoaRays io{sc.nodes, sc.tris, n, rays, t, u, v, prim, tie, nullptr};
  const int64_t nchunks = ((int64_t)n + chunk - 1) / chunk;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const int begin = (int)(c * chunk);
    trace_engine<false, RING>(io, begin, (int)min((int64_t)begin + chunk, (int64_t)n), stk, &nv, &nt, &ovf);
  }
  if (ovf) atomicAdd(overflow, ovf);
}

// ---------------------------------------------------------------- launchers
// LDS ring variants: RING x 128 lanes x 4 B of LDS per block.
#define ORION_RING_DISPATCH(S, CALL)                     \
  do {                                                   \
    if ((S) <= 8) { constexpr int ST = 8; CALL; }        \
    else if ((S) <= 16) { constexpr int ST = 16; CALL; } \
    else { constexpr int ST = 32; CALL; }                \
  } while (0)

int ring_variant(int ring) {
  if (ring <= 0) return kDefaultRing;
  return ring <= 8 ? 8 : ring <= 16 ? 16 : 32;
}

int trace_waves_per_simd(int ring) {
  ring = ring_variant(ring);
  int nb = 0;
  hipError_t e = hipSuccess;
  ORION_RING_DISPATCH(ring, e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<false, ST, false>, kTraceBlock, 0));
  return e == hipSuccess ? nb * (kTraceBlock / 64) / 4 : 0;
}

size_t trace_spill_ints(int stack_bound, int ring, int blocks) {
  int cap = std::max(0, stack_bound - ring_variant(ring));
  return (size_t)cap * blocks * kTraceBlock;
}

int trace_chunk(int chunk) { return chunk > 0 ? std::min(std::max(chunk, 64), 4096) : kDefaultChunk; }

int trace_grid(int nseg, int seg, bool any, int chunk) {
  chunk = trace_chunk(chunk);
  const int64_t seg_entries = any ? 2 * (int64_t)seg : seg;
  const int64_t waves = nseg * ((seg_entries + chunk - 1) / chunk);
  return (int)std::max<int64_t>(1, (waves + kTraceBlock / 64 - 1) / (kTraceBlock / 64));
}

void launch_raygen(const Job *jobs, const Wave &w, bool multi, hipStream_t st) {
  dim3 g(((size_t)w.nseg * w.seg + 255) / 256);
  if (multi) hipLaunchKernelGGL(k_raygen<true>, g, dim3(256), 0, st, jobs, w);
  else hipLaunchKernelGGL(k_raygen<false>, g, dim3(256), 0, st, jobs, w);
}
void launch_trace_pair(const Job *jobs, const Wave &w, int ring, int chunk, bool multi, hipStream_t st, int cps_c,
                       int cps_a, int rec) {
  ring = ring_variant(ring);
  chunk = trace_chunk(chunk);
  const int full_c = (w.seg + chunk - 1) / chunk, full_a = (int)((2 * (int64_t)w.seg + chunk - 1) / chunk);
  if (cps_c <= 0 || cps_c > full_c) cps_c = full_c;
  if (cps_a <= 0 || cps_a > full_a) cps_a = full_a;
  const int nc = (std::max(1, w.nseg * cps_c) + kXcds - 1) / kXcds * kXcds;
  const dim3 g(nc + std::max(1, w.nseg * cps_a));
  if (multi)
    ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace_pair<ST, true>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk,
                                                 nc, cps_c, cps_a, rec));
  else
    ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace_pair<ST, false>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk,
                                                 nc, cps_c, cps_a, rec));
}

int trace_pair_grid(int nseg, int seg, int chunk) {
  chunk = trace_chunk(chunk);
  return (trace_grid(nseg, seg, false, chunk) + kXcds - 1) / kXcds * kXcds + trace_grid(nseg, seg, true, chunk);
}

void launch_trace(const Job *jobs, const Wave &w, bool any, int ring, int chunk, bool multi, hipStream_t st, int cps,
                  int rec) {
  ring = ring_variant(ring);
  chunk = trace_chunk(chunk);
  const int full = (int)(((any ? 2 * (int64_t)w.seg : w.seg) + chunk - 1) / chunk);
  if (cps <= 0 || cps > full) cps = full;
  const dim3 g((unsigned)std::max(1, w.nseg * cps));
  if (multi) {
    if (any) ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace<true, ST, true>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk, cps, rec));
    else ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace<false, ST, true>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk, cps, rec));
  } else {
    if (any) ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace<true, ST, false>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk, cps, rec));
    else ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace<false, ST, false>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk, cps, rec));
  }
}
void launch_trace_photons(const Job *jobs, const Wave &w, int ring, int chunk, hipStream_t st, int cps, int rec) {
  ring = ring_variant(ring);
  chunk = trace_chunk(chunk);
  const int full = (w.seg + chunk - 1) / chunk;
  if (cps <= 0 || cps > full) cps = full;
  const dim3 g((unsigned)std::max(1, w.nseg * cps));
  ORION_RING_DISPATCH(ring, hipLaunchKernelGGL((k_trace_photons<ST>), g, dim3(kTraceBlock), 0, st, jobs, w, chunk, cps, rec));
}
void launch_between(const Wave &w, int nxt, hipStream_t st) {
  hipLaunchKernelGGL(k_between, dim3(1), dim3(64), 0, st, w, nxt);
}
void launch_shade(const Job *jobs, const Wave &w, int bounce, bool multi, hipStream_t st) {
  const int blocks = w.nseg * w.shade_parts;
  if (multi) hipLaunchKernelGGL(k_shade<true>, dim3(blocks), dim3(kShadeThreads), 0, st, jobs, w, bounce);
  else hipLaunchKernelGGL(k_shade<false>, dim3(blocks), dim3(kShadeThreads), 0, st, jobs, w, bounce);
}
void launch_film(const Job *jobs, const Wave &w, int max_pix, bool multi, hipStream_t st) {
  int blocks = (max_pix + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (multi) hipLaunchKernelGGL(k_film<true>, dim3(blocks), dim3(256), 0, st, jobs, w);
  else hipLaunchKernelGGL(k_film<false>, dim3(blocks), dim3(256), 0, st, jobs, w);
}
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, int *spill, int spill_cap,
                            unsigned long long *stats, hipStream_t st) {
  constexpr int ST = kDefaultRing;
  const int chunk = kDefaultChunk;
  const dim3 g(trace_grid(1, n, false, chunk));
  if (any)
    hipLaunchKernelGGL((k_intersect_batch<true, ST>), g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie,
                       occluded, spill, spill_cap, chunk, stats);
  else
    hipLaunchKernelGGL((k_intersect_batch<false, ST>), g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie,
                       occluded, spill, spill_cap, chunk, stats);
}
void launch_intersect_fixed(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *spill, int spill_cap, int blocks, unsigned int *overflow,
                            hipStream_t st) {
  constexpr int ST = kDefaultRing;
  hipLaunchKernelGGL((k_intersect_fixed<ST>), dim3(blocks), dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v,
                     tie, spill, spill_cap, kDefaultChunk, overflow);
}

}  // namespace orion_amd
