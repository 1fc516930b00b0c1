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
#include "wavefront.h"

namespace orion_amd {
using namespace dev;

// ---------------------------------------------------------------- camera
// pinhole.cpp:8-33 with next2() = (x: dim d+1, y: dim d), sampler.h:31 as
// compiled by g++ (the right argument is evaluated first).
__device__ __forceinline__ void camera_ray(const DScene &sc, int i, int j, uint32_t index, V3 *o, V3 *d) {
  float uy = sobol(sc.sobol, index, 5);
  float ux = sobol(sc.sobol, index, 6);
  float rx = (float)i + ux, ry = (float)j + uy;
  const float *c = sc.cam;
  float width = c[14], height = c[15];
  float xx = (2 * rx / width - 1) * c[13] * c[12];
  float yy = (1 - 2 * ry / height) * c[12];
  V3 dir = normalize(mk(xx, yy, -1.0f));
  *o = mk(c[0], c[1], c[2]);
  *d = mat3(c + 3, dir);
}

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
    atomicAdd(&w.stats[S_NODES], (unsigned long long)nv);
    atomicAdd(&w.stats[S_TRIS], (unsigned long long)nt);
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

// ---------------------------------------------------------------- lights
// Distribution1D::sampleDiscrete (sampling.h:110-120) over light power.
__device__ __forceinline__ int sample_light(const DScene &sc, float u, float *pdf) {
  int size = sc.n_lights + 1;
  int first = 0, len = size;
  while (len > 0) {
    int half = len >> 1, mid = first + half;
    if (sc.light_cdf[mid] <= u) { first = mid + 1; len -= half + 1; }
    else len = half;
  }
  int off = first - 1;
  off = off < 0 ? 0 : (off > size - 2 ? size - 2 : off);
  *pdf = (sc.light_func_int > 0) ? sc.light_func[off] / (sc.light_func_int * sc.n_lights) : 0;
  return off;
}

__device__ __forceinline__ float spot_falloff(const float *L, V3 w) {  // spot.cpp:10-19
  V3 wl = normalize(mat3(L + 11, w));
  float cosTheta = wl.z;
  if (cosTheta < L[9]) return 0;
  if (cosTheta > L[10]) return 1;
  float delta = (cosTheta - L[9]) / (L[10] - L[9]);
  return (delta * delta) * (delta * delta);
}

// Light::sample_Li; *so / *sd / *stmax = the ShadowTester ray (intersection.h:40-44)
__device__ inline Sp light_sample_li(const DScene &sc, int l, V3 refp, V3 refng, float u0, float u1, V3 *wi,
                                     float *pdf, V3 *so, V3 *sd) {
  const float *L = sc.lights + 32 * (size_t)l;
  int type = (int)L[0];
  Sp Le = sp3(L[2], L[3], L[4]);
  V3 target;
  if (type == 0) {  // diffuse.cpp:13-25 + triangle.cpp:85-121
    V3 p0, p1, p2, n0, n1, n2;
    int m_, l_;
    load_prim(sc, (int)L[1], &p0, &p1, &p2, &n0, &n1, &n2, &m_, &l_);
    float su0 = sqrtf(u0);
    float b0 = 1 - su0, b1 = u1 * su0;
    V3 p = add(add(mul(p0, b0), mul(p1, b1)), mul(p2, 1 - b0 - b1));
    V3 g = normalize(cross(sub(p1, p0), sub(p2, p0)));
    V3 ns = add(add(mul(n0, b0), mul(n1, b1)), mul(n2, 1.0f - b0 - b1));
    V3 ng = dot(g, ns) < 0.f ? neg(g) : g;
    float pd = 1.0f / sc.light_area[l];
    V3 w = sub(p, refp);
    if (lensq(w) == 0) pd = 0;
    else {
      w = normalize(w);
      pd *= lensq(sub(refp, p)) / absdot(ng, neg(w));
      if (is_inf(pd)) pd = 0;
    }
    if (pd == 0 || lensq(sub(p, refp)) == 0) { *pdf = 0; return sp(0); }
    *pdf = pd;
    *wi = normalize(sub(p, refp));
    target = p;
    Le = dot(ng, neg(*wi)) > 0 ? Le : sp(0);
  } else {  // point.cpp:10-19, spot.cpp:20-29
    V3 pos = mk(L[5], L[6], L[7]);
    *wi = normalize(sub(pos, refp));
    *pdf = 1.0f;
    target = pos;
    if (type == 1) Le = Le / lensq(sub(pos, refp));
    else Le = Le * spot_falloff(L, neg(*wi)) / lensq(sub(pos, refp));
  }
  *so = offset_ray_origin(refp, refng, sub(target, refp));
  *sd = sub(target, *so);
  return Le;
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
        Sp Ld = sp(0);
        if ((fl & F_SHADOW) && w.vis[p]) acc(Ld, sp3(w.Ar[p], w.Ag[p], w.Ab[p]));
        if (fl & F_MIS) {
          int hp = w.hprim[N + p];
          int l = w.nlight[p];
          const float *LR = sc.lights + 32 * (size_t)l;
          if (hp >= 0 && hp == (int)LR[1]) {
            // Triangle::pdf (triangle.cpp:123-135) on the MIS ray's hit
            Isect in;
            make_isect(sc, hp, w.hu[N + p], w.hv[N + p], &in);
            V3 wi = mk(w.dx[N + p], w.dy[N + p], w.dz[N + p]);
            V3 pp = mk(w.px[p], w.py[p], w.pz[p]);
            float lpdf = lensq(sub(pp, in.p)) / (absdot(in.ng, neg(wi)) * sc.light_area[l]);
            if (is_inf(lpdf)) lpdf = 0;
            if (lpdf != 0) {
              float bp = w.bpdf[p];
              float weight = power_heuristic(bp, lpdf);
              Sp Lm = dot(in.ng, neg(wi)) > 0 ? sp3(LR[2], LR[3], LR[4]) : sp(0);
              if (!black(Lm)) acc(Ld, sp3(w.fr[p], w.fg[p], w.fb[p]) * Lm * weight / bp);
            }
          }
        }
        acc(L, sp3(w.nbr[p], w.nbg[p], w.nbb[p]) * (Ld / w.sel[p]));
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
        if ((b == 0 || (fl & F_SPEC)) && hp >= 0 && is.light >= 0) {
          const float *LR = sc.lights + 32 * (size_t)is.light;
          Sp Le = dot(is.ng, neg(d)) > 0 ? sp3(LR[2], LR[3], LR[4]) : sp(0);  // diffuse.cpp:53-56
          acc(L, beta * Le);
        }
        if (hp < 0 || b >= maxDepth) {
          fl &= ~F_ALIVE;
        } else {
          Bsdf bs;
          make_bsdf(sc.mats + 16 * (size_t)is.mat, is.ng, is.ns, is.u, is.v, false, &bs);
          uint32_t index = path_index(w, p);
          int dim = w.dim[p];
          const int nonspec = ALL & ~SPEC;
          if (bsdf_num(bs, nonspec) > 0 && sc.n_lights > 0) {
            float sel;
            int l = sample_light(sc, sobol(sc.sobol, index, dim++), &sel);
            if (sel != 0) {
              float lu1 = sobol(sc.sobol, index, dim++), lu0 = sobol(sc.sobol, index, dim++);
              float bu1 = sobol(sc.sobol, index, dim++), bu0 = sobol(sc.sobol, index, dim++);
              V3 wo = normalize(neg(d));
              V3 wi = mk(0, 0, 0), so, sdir;
              float lpdf = 0;
              Sp Li = light_sample_li(sc, l, is.p, is.ng, lu0, lu1, &wi, &lpdf, &so, &sdir);
              bool delta = (int)sc.lights[32 * (size_t)l] != 0;
              bool hs = false, hm = false;
              if (lpdf > 0 && !black(Li)) {
                Sp f = bsdf_f(bs, wi, wo, nonspec) * absdot(wi, is.ns);
                float bpdf = bsdf_pdf(bs, wi, wo, nonspec);
                if (!black(f)) {
                  Sp A;
                  if (delta) A = f * Li / lpdf;
                  else A = f * Li * power_heuristic(lpdf, bpdf) / lpdf;
                  w.Ar[p] = A.r; w.Ag[p] = A.g; w.Ab[p] = A.b;
                  int s2 = 2 * N + p;
                  w.ox[s2] = so.x; w.oy[s2] = so.y; w.oz[s2] = so.z;
                  w.dx[s2] = sdir.x; w.dy[s2] = sdir.y; w.dz[s2] = sdir.z;
                  w.tmax[s2] = 1 - kShadowEps;
                  hs = true;
                }
              }
              if (!delta) {
                float bp2 = 0;
                int st = 0;
                V3 wi2 = wi;
                Sp f2 = bsdf_sample(bs, &wi2, wo, bu0, bu1, &bp2, nonspec, &st);
                mulby(f2, sp(absdot(wi2, is.ns)));
                if (!black(f2) && bp2 > 0) {
                  V3 o2 = offset_ray_origin(is.p, is.ng, wi2);
                  int s1 = N + p;
                  w.ox[s1] = o2.x; w.oy[s1] = o2.y; w.oz[s1] = o2.z;
                  w.dx[s1] = wi2.x; w.dy[s1] = wi2.y; w.dz[s1] = wi2.z;
                  w.tmax[s1] = kFltMax;
                  w.fr[p] = f2.r; w.fg[p] = f2.g; w.fb[p] = f2.b;
                  w.bpdf[p] = bp2;
                  w.px[p] = is.p.x; w.py[p] = is.p.y; w.pz[p] = is.p.z;
                  w.nlight[p] = l;
                  hm = true;
                }
              }
              if (hs || hm) {
                w.sel[p] = sel;
                w.nbr[p] = beta.r; w.nbg[p] = beta.g; w.nbb[p] = beta.b;
                fl |= F_PEND | (hs ? F_SHADOW : 0) | (hm ? F_MIS : 0);
                want_shadow = hs;
                want_ray1 = hm;
              }
            }
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
