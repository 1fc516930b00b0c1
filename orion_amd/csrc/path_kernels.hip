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
#include <hip/hip_runtime.h>

#include "device_scene.h"
#include "kernels.h"
#include "shading.h"
#include "wavefront.h"

namespace orion_amd {
using namespace dev;

__global__ void k_raygen(DScene sc, Wave w) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    w.ctrl[C_PATH0] = w.N;
    w.ctrl[C_RAY] = w.N;
    w.ctrl[C_SHADOW] = 0;
    w.ctrl[C_WORK_RAY] = 0;
    w.ctrl[C_WORK_SHADOW] = 0;
  }
  if (p >= w.N) return;
  int s = p / w.npix, pix = p - s * w.npix;
  int j = pix / w.W, i = pix - j * w.W;
  uint32_t index = (w.k0 + (uint32_t)s) * (uint32_t)w.npix + (uint32_t)pix;
  V3 o, d;
  camera_ray(sc, i, j, index, &o, &d);
  w.ox[p] = o.x; w.oy[p] = o.y; w.oz[p] = o.z;
  w.dx[p] = d.x; w.dy[p] = d.y; w.dz[p] = d.z;
  w.tmax[p] = kFltMax;
  w.br[p] = 1.f; w.bg[p] = 1.f; w.bb[p] = 1.f;
  w.Lr[p] = 0.f; w.Lg[p] = 0.f; w.Lb[p] = 0.f;
  w.dim[p] = 7;
  w.bounce[p] = 0;
  w.flags[p] = F_ALIVE;
  w.qpath[0][p] = p;
  w.qray[p] = p;
}

// ---------------------------------------------------------------- traversal
// Persistent: each wave pulls 64-ray tiles from the queue until it drains.
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace(DScene sc, Wave w) {
  __shared__ int stack[kStackEntries * kTraceBlock];
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

__device__ __forceinline__ uint32_t path_index(const Wave &w, int p) {
  int s = p / w.npix;
  return (w.k0 + (uint32_t)s) * (uint32_t)w.npix + (uint32_t)(p - s * w.npix);
}

// ---------------------------------------------------------------- shade
__global__ __launch_bounds__(256) void k_shade(DScene sc, Wave w, int cur, int maxDepth) {
  const int nxt = cur ^ 1;
  const int count = w.ctrl[C_PATH0 + cur];
  const int N = w.N;
  for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    int i = base + threadIdx.x;
    bool act = i < count;
    int p = act ? w.qpath[cur][i] : 0;
    bool want_ray0 = false, want_ray1 = false, want_shadow = false, keep = false, done = false;
    if (act) {
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
        if (hp < 0 || b >= maxDepth) {
          fl &= ~F_ALIVE;
        } else {
          Bsdf bs;
          make_bsdf(sc.mats + 16 * (size_t)is.mat, is.ng, is.ns, is.u, is.v, false, &bs);
          uint32_t index = path_index(w, p);
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
              float q = mx(0.05f, 1 - maxc(beta));
              if (sobol(sc.sobol, index, dim++) < q) alive = false;
              else beta = beta / (1 - q);
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
__global__ void k_film(Wave w, int S, float *film, int *count) {
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= w.npix) return;
  float r = film[3 * pix], g = film[3 * pix + 1], b = film[3 * pix + 2];
  for (int s = 0; s < S; ++s) {
    int p = s * w.npix + pix;
    r += w.Lr[p];
    g += w.Lg[p];
    b += w.Lb[p];
  }
  film[3 * pix] = r;
  film[3 * pix + 1] = g;
  film[3 * pix + 2] = b;
  count[pix] += S;
}

// ---------------------------------------------------------------- batch intersection API
template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_intersect_batch(DScene sc, int n, const float *rays, float *t,
                                                                 int *prim, float *u, float *v, int *tie,
                                                                 int *occluded) {
  __shared__ int stack[kStackEntries * kTraceBlock];
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
void launch_raygen(const DScene &sc, const Wave &w, hipStream_t st) {
  hipLaunchKernelGGL(k_raygen, dim3((w.N + 255) / 256), dim3(256), 0, st, sc, w);
}
void launch_trace(const DScene &sc, const Wave &w, bool any, int blocks, hipStream_t st) {
  if (any) hipLaunchKernelGGL(k_trace<true>, dim3(blocks), dim3(kTraceBlock), 0, st, sc, w);
  else hipLaunchKernelGGL(k_trace<false>, dim3(blocks), dim3(kTraceBlock), 0, st, sc, w);
}
void launch_between(const Wave &w, int nxt, hipStream_t st) {
  hipLaunchKernelGGL(k_between, dim3(1), dim3(64), 0, st, w, nxt);
}
void launch_shade(const DScene &sc, const Wave &w, int cur, int maxDepth, hipStream_t st) {
  int blocks = (w.N + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, st, sc, w, cur, maxDepth);
}
void launch_film(const Wave &w, int S, float *film, int *count, hipStream_t st) {
  hipLaunchKernelGGL(k_film, dim3((w.npix + 255) / 256), dim3(256), 0, st, w, S, film, count);
}
void launch_intersect_batch(const DScene &sc, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occluded, bool any, hipStream_t st) {
  dim3 g((n + kTraceBlock - 1) / kTraceBlock);
  if (any) hipLaunchKernelGGL(k_intersect_batch<true>, g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie, occluded);
  else hipLaunchKernelGGL(k_intersect_batch<false>, g, dim3(kTraceBlock), 0, st, sc, n, rays, t, prim, u, v, tie, occluded);
}

}  // namespace orion_amd
