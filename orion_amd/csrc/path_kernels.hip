// Wavefront path tracer for gfx950 (PathTracing::Li, integrator/path.cpp:50-99,
// with direct lighting from core/integrator.cpp:8-90).
//
// Per wave: raygen -> [closest-hit trace (continuation + MIS rays) ->
// any-hit trace (shadow rays) -> shade] x (maxDepth + 1) -> film.  The shade
// kernel first resolves the direct-lighting estimate pending from the
// previous bounce (so L accumulates in the reference's order: emission of
// bounce b, direct light of bounce b, emission of bounce b + 1, ...), then
// shades the new hit: emission, light sampling, BSDF sampling, Russian
// roulette.  Queues are compacted with wave64 ballots (one atomic per wave).
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

__device__ __forceinline__ int find_job(const Job *jobs, int njobs, int p) {
  int lo = 0, hi = njobs - 1;  // last job with path0 <= p
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].path0 <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool MULTI>
__global__ void k_raygen(const Job *jobs, Wave w) {
  const int npaths = w.ctrl[C_NPATHS];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    w.ctrl[C_PATH0] = npaths;
    w.ctrl[C_RAY] = npaths;
    w.ctrl[C_SHADOW] = 0;
    w.ctrl[C_WORK_RAY] = 0;
    w.ctrl[C_WORK_SHADOW] = 0;
  }
  if (p >= npaths) return;
  int jb = MULTI ? find_job(jobs, w.ctrl[C_NJOBS], p) : 0;
  const Job &J = jobs[jb];
  int q = p - J.path0;
  int s = q / J.npix, pix = q - s * J.npix;
  int j = pix / J.W, i = pix - j * J.W;
  uint32_t index = (J.k0 + (uint32_t)s) * (uint32_t)J.npix + (uint32_t)pix;  // path.cpp:26
  V3 o, d;
  camera_ray(J.sc, i, j, index, &o, &d);
  w.ox[p] = o.x; w.oy[p] = o.y; w.oz[p] = o.z;
  w.dx[p] = d.x; w.dy[p] = d.y; w.dz[p] = d.z;
  w.tmax[p] = kFltMax;
  w.br[p] = 1.f; w.bg[p] = 1.f; w.bb[p] = 1.f;
  w.Lr[p] = 0.f; w.Lg[p] = 0.f; w.Lb[p] = 0.f;
  w.dim[p] = 7;
  w.bounce[p] = 0;
  w.flags[p] = F_ALIVE;
  if (MULTI) w.job[p] = jb;
  w.qpath[0][p] = p;
  w.qray[p] = p;
}

// ---------------------------------------------------------------- traversal
// Persistent: each wave pulls 64-ray tiles from the queue until it drains.
template <bool ANY, int STACK, bool MULTI>
__global__ __launch_bounds__(kTraceBlock) void k_trace(const Job *jobs, Wave w) {
  __shared__ int stack[STACK * kTraceBlock];
  int *stk = stack + threadIdx.x;
  const int count = w.ctrl[ANY ? C_SHADOW : C_RAY];
  int *work = w.ctrl + (ANY ? C_WORK_SHADOW : C_WORK_RAY);
  const int *queue = ANY ? w.qshadow : w.qray;
  int lane = lane_id();
  uint32_t nv = 0, nt = 0;
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(work, 64);
    base = __shfl(base, 0);
    if (base >= count) break;
    int q = base + lane;
    if (q < count) {
      int r = queue[q];
      int slot = ANY ? 2 * w.N + r : r;  // shadow rays live in slot 2
      int p = ANY ? r : (r >= w.N ? r - w.N : r);
      const DScene &sc = jobs[MULTI ? w.job[p] : 0].sc;
      V3 o = mk(w.ox[slot], w.oy[slot], w.oz[slot]);
      V3 d = mk(w.dx[slot], w.dy[slot], w.dz[slot]);
      HitRec h;
      bool hit = traverse<ANY>(sc, o, d, 0.0f, w.tmax[slot], &h, stk, kTraceBlock, &nv, &nt);
      if (ANY) {
        w.vis[r] = hit ? 0 : 1;
      } else {
        w.hprim[r] = hit ? h.prim : -1;
        w.ht[r] = h.t;
        w.hu[r] = h.u;
        w.hv[r] = h.v;
      }
    }
  }
  // per-wave reduction of traversal counters (for the roofline's n_node / n_tri)
  for (int off = 32; off > 0; off >>= 1) {
    nv += __shfl_down(nv, off);
    nt += __shfl_down(nt, off);
  }
  if (lane == 0 && w.stats) {
    atomicAdd(&w.stats[ANY ? S_NODES_S : S_NODES_C], (unsigned long long)nv);
    atomicAdd(&w.stats[ANY ? S_TRIS_S : S_TRIS_C], (unsigned long long)nt);
  }
}

__global__ void k_between(Wave w, int nxt) {
  // runs after both traces, before shade: account the rays just traced and
  // reset the queues the shade kernel appends to
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (w.stats) {
      w.stats[S_CLOSEST] += (unsigned long long)w.ctrl[C_RAY];
      w.stats[S_SHADOW] += (unsigned long long)w.ctrl[C_SHADOW];
    }
    w.ctrl[C_RAY] = 0;
    w.ctrl[C_SHADOW] = 0;
    w.ctrl[C_WORK_RAY] = 0;
    w.ctrl[C_WORK_SHADOW] = 0;
    w.ctrl[C_PATH0 + nxt] = 0;
  }
}

// ---------------------------------------------------------------- shade
template <bool MULTI>
__global__ __launch_bounds__(256) void k_shade(const Job *jobs, Wave w, int cur) {
  const int nxt = cur ^ 1;
  const int count = w.ctrl[C_PATH0 + cur];
  const int N = w.N;
  for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    int i = base + threadIdx.x;
    bool act = i < count;
    int p = act ? w.qpath[cur][i] : 0;
    bool want_ray0 = false, want_ray1 = false, want_shadow = false, keep = false, done = false;
    if (act) {
      const Job &J = jobs[MULTI ? w.job[p] : 0];
      const DScene &sc = J.sc;
      int fl = w.flags[p];
      Sp L = sp3(w.Lr[p], w.Lg[p], w.Lb[p]);
      // 1) resolve the direct-lighting estimate of the previous bounce
      if (fl & F_PEND) {
        resolve_direct(sc, w, p, fl, &L);
        fl &= ~(F_PEND | F_SHADOW | F_MIS);
      }
      // 2) the hit of this bounce
      if (fl & F_ALIVE) {
        int b = w.bounce[p];
        int hp = w.hprim[p];
        Sp beta = sp3(w.br[p], w.bg[p], w.bb[p]);
        V3 d = mk(w.dx[p], w.dy[p], w.dz[p]);
        Isect is;
        if (hp >= 0) make_isect(sc, hp, w.hu[p], w.hv[p], &is);
        if ((b == 0 || (fl & F_SPEC)) && hp >= 0 && is.light >= 0) acc(L, beta * area_emission(sc, is, neg(d)));
        if (hp < 0 || b >= J.maxDepth) {
          fl &= ~F_ALIVE;
        } else {
          Bsdf bs;
          make_bsdf(sc.mats + 16 * (size_t)is.mat, is.ng, is.ns, is.u, is.v, false, &bs);
          int q = p - J.path0;
          int s = q / J.npix;
          uint32_t index = (J.k0 + (uint32_t)s) * (uint32_t)J.npix + (uint32_t)(q - s * J.npix);
          int dim = w.dim[p];
          if (bsdf_num(bs, ALL & ~SPEC) > 0) {
            int nf = setup_direct(sc, w, p, bs, is, d, beta, index, &dim, false);
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
              w.ox[p] = o.x; w.oy[p] = o.y; w.oz[p] = o.z;
              w.dx[p] = wi.x; w.dy[p] = wi.y; w.dz[p] = wi.z;
              w.tmax[p] = kFltMax;
              w.bounce[p] = b + 1;
              want_ray0 = true;
            } else {
              fl &= ~F_ALIVE;
            }
          }
          w.br[p] = beta.r; w.bg[p] = beta.g; w.bb[p] = beta.b;
          w.dim[p] = dim;
        }
      }
      w.Lr[p] = L.r; w.Lg[p] = L.g; w.Lb[p] = L.b;
      w.flags[p] = fl;
      keep = (fl & (F_ALIVE | F_PEND)) != 0;
      done = !keep;
    }
    int a = wave_append(&w.ctrl[C_PATH0 + nxt], keep);
    if (keep) w.qpath[nxt][a] = p;
    a = wave_append(&w.ctrl[C_RAY], want_ray0);
    if (want_ray0) w.qray[a] = p;
    a = wave_append(&w.ctrl[C_RAY], want_ray1);
    if (want_ray1) w.qray[a] = N + p;
    a = wave_append(&w.ctrl[C_SHADOW], want_shadow);
    if (want_shadow) w.qshadow[a] = p;
    unsigned long long dm = __ballot(done);
    unsigned long long mm = __ballot(want_ray1);
    if (lane_id() == 0 && w.stats) {
      if (dm) atomicAdd(&w.stats[S_PATHS], (unsigned long long)__popcll(dm));
      if (mm) atomicAdd(&w.stats[S_REFEQ], (unsigned long long)__popcll(mm));
    }
  }
}

// RenderTarget::acumulate in sample order per pixel (rendertarget.cpp:22-27)
template <bool MULTI>
__global__ void k_film(const Job *jobs, Wave w) {
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
      r += w.Lr[p];
      gg += w.Lg[p];
      b += w.Lb[p];
    }
    J.film[3 * pix] = r;
    J.film[3 * pix + 1] = gg;
    J.film[3 * pix + 2] = b;
    J.count[pix] += J.S;
  }
}

// ---------------------------------------------------------------- batch intersection API
template <bool ANY, int STACK>
__global__ __launch_bounds__(kTraceBlock) void k_intersect_batch(DScene sc, int n, const float *rays, float *t,
                                                                 int *prim, float *u, float *v, int *tie,
                                                                 int *occluded) {
  __shared__ int stack[STACK * kTraceBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 o = mk(rays[i], rays[n + i], rays[2 * n + i]);
  V3 d = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
  float tmin = rays[6 * n + i], tmx = rays[7 * n + i];
  HitRec h;
  bool hit = traverse<ANY>(sc, o, d, tmin, tmx, &h, stack + threadIdx.x, kTraceBlock, nullptr, nullptr);
  if (ANY) {
    occluded[i] = hit ? 1 : 0;
  } else {
    t[i] = hit ? h.t : 0.f;
    prim[i] = hit ? h.prim : -1;
    if (u) u[i] = hit ? h.u : 0.f;
    if (v) v[i] = hit ? h.v : 0.f;
    if (tie) tie[i] = hit ? h.tie : 0;
  }
}

// ---------------------------------------------------------------- launchers
// Stack-depth variants: the smallest that holds the BVH depth + 2, so LDS
// (STACK x 128 lanes x 4 B per block) caps occupancy no more than needed.
#define ORION_STACK_DISPATCH(S, CALL)                     \
  do {                                                    \
    if ((S) <= 16) { constexpr int ST = 16; CALL; }       \
    else if ((S) <= 24) { constexpr int ST = 24; CALL; }  \
    else if ((S) <= 32) { constexpr int ST = 32; CALL; }  \
    else { constexpr int ST = kStackEntries; CALL; }      \
  } while (0)

int stack_variant(int depth) {
  int need = depth + 2;
  return need <= 16 ? 16 : need <= 24 ? 24 : need <= 32 ? 32 : kStackEntries;
}

int trace_blocks_per_cu(int stack_entries, bool any) {
  static int cache[2][4] = {{0}};
  int vi = stack_entries <= 16 ? 0 : stack_entries <= 24 ? 1 : stack_entries <= 32 ? 2 : 3;
  int &c = cache[any ? 1 : 0][vi];
  if (c == 0) {
    int nb = 0;
    hipError_t e = hipSuccess;
    if (any) ORION_STACK_DISPATCH(stack_entries, e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<true, ST, true>, kTraceBlock, 0));
    else ORION_STACK_DISPATCH(stack_entries, e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<false, ST, true>, kTraceBlock, 0));
    c = (e == hipSuccess && nb > 0) ? nb : 4;
  }
  return c;
}

void launch_raygen(const Job *jobs, const Wave &w, int max_paths, bool multi, hipStream_t st) {
  dim3 g((max_paths + 255) / 256);
  if (multi) hipLaunchKernelGGL(k_raygen<true>, g, dim3(256), 0, st, jobs, w);
  else hipLaunchKernelGGL(k_raygen<false>, g, dim3(256), 0, st, jobs, w);
}
void launch_trace(const Job *jobs, const Wave &w, bool any, int stack_entries, int blocks, bool multi,
                  hipStream_t st) {
  if (multi) {
    if (any) ORION_STACK_DISPATCH(stack_entries, hipLaunchKernelGGL((k_trace<true, ST, true>), dim3(blocks), dim3(kTraceBlock), 0, st, jobs, w));
    else ORION_STACK_DISPATCH(stack_entries, hipLaunchKernelGGL((k_trace<false, ST, true>), dim3(blocks), dim3(kTraceBlock), 0, st, jobs, w));
  } else {
    if (any) ORION_STACK_DISPATCH(stack_entries, hipLaunchKernelGGL((k_trace<true, ST, false>), dim3(blocks), dim3(kTraceBlock), 0, st, jobs, w));
    else ORION_STACK_DISPATCH(stack_entries, hipLaunchKernelGGL((k_trace<false, ST, false>), dim3(blocks), dim3(kTraceBlock), 0, st, jobs, w));
  }
}
void launch_between(const Wave &w, int nxt, hipStream_t st) {
  hipLaunchKernelGGL(k_between, dim3(1), dim3(64), 0, st, w, nxt);
}
void launch_shade(const Job *jobs, const Wave &w, int cur, int max_paths, bool multi, hipStream_t st) {
  int blocks = (max_paths + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (multi) hipLaunchKernelGGL(k_shade<true>, dim3(blocks), dim3(256), 0, st, jobs, w, cur);
  else hipLaunchKernelGGL(k_shade<false>, dim3(blocks), dim3(256), 0, st, jobs, w, cur);
}
void launch_film(const Job *jobs, const Wave &w, int max_pix, bool multi, hipStream_t st) {
  int blocks = (max_pix + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (multi) hipLaunchKernelGGL(k_film<true>, dim3(blocks), dim3(256), 0, st, jobs, w);
  else hipLaunchKernelGGL(k_film<false>, dim3(blocks), dim3(256), 0, st, jobs, w);
}
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, hipStream_t st) {
  dim3 g((n + kTraceBlock - 1) / kTraceBlock);
  if (any) ORION_STACK_DISPATCH(sc.stack_entries, hipLaunchKernelGGL((k_intersect_batch<true, ST>), g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie, occluded));
  else ORION_STACK_DISPATCH(sc.stack_entries, hipLaunchKernelGGL((k_intersect_batch<false, ST>), g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie, occluded));
}

}  // namespace orion_amd
