// SPPM (SPPM::render, integrator/sppm.cpp:58-322) on gfx950.
//
// Per iteration:
//   camera pass   wavefront over pixels (closest + shadow + MIS rays, the
//                 deferred direct lighting of the path tracer with a uniform
//                 light pick), storing a visible point at the first diffuse
//                 hit (or glossy hit at the last depth)
//   grid          bounds / resolution reduction, per visible point the cells
//                 its radius overlaps, (hash key, pixel) pairs, a stable LSD
//                 radix sort of the keys (8-bit digits, wave64 ballot ranking),
//                 per-cell [start, end) ranges
//   photon pass   wavefront over photons: light emission, closest hits,
//                 deposits into the visible points of the hashed cell (float
//                 atomics for Phi, integer atomics for M and per-cell counts),
//                 BSDF sampling in Photon mode, Russian roulette
//   update        radius / N / tau update, then the per-iteration film add
// The hash keeps the reference's collisions: a visible point overlapping two
// cells that hash alike is listed (and deposited into) twice, as with orion's
// per-cell linked lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"
#include "device_scene.h"
#include "host_scene.h"
#include "kernels.h"
#include "runtime.h"
#include "shading.h"

namespace orion_amd {
using namespace dev;

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
      return ORION_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

struct Pix {  // per-pixel state, SoA
  float *radius, *N, *Ldr, *Ldg, *Ldb, *taur, *taug, *taub, *Phir, *Phig, *Phib;
  int *M;
  // visible point
  float *vpx, *vpy, *vpz, *vwox, *vwoy, *vwoz;
  float *vng, *vns, *vu, *vv;  // 3 floats per pixel each (xyz interleaved)
  int *vmat;
  float *vbr, *vbg, *vbb;
};

struct Grid {
  float *fctl;      // [0..2] lo, [3..5] hi, [6] maxRadius
  int *ictl;        // [0..2] gridRes, [3] any, [4..9] sortable lo/hi, [10] maxRadius bits, [11] n_pairs
  int *cnt;         // per pixel pair counts -> exclusive offsets (npix + 1)
  uint32_t *keys[2];
  int *vals[2];
  int *cell_start, *cell_end, *photon_cell;
  int *part;        // the bounds kernel's per-block partials (8 ints per block, kBoundsBlocks blocks)
  int *hist;        // radix histogram (256 x tiles) -> offsets
  int *scan_tmp;
  int hash_size;
  int64_t pair_cap = 0;  // entries of keys / vals
};

// The camera passes of a batch of iterations run on their own stream beside
// the previous batch's grids, deposits and photons (they depend on nothing
// those touch), so they have their own wave buffers and traversal spill
// area.  Iteration i writes visible-point set i % nsets (px[i % nsets]: the
// pixel state is shared, each set has its own visible-point fields, set s
// at s x npix in field-major arrays), and the film reads the snapshot of Ld
// the replay takes after iteration i (set i % nsets of ld_snap).  nsets is
// two batches' worth, so one batch's camera passes fill the sets the
// previous batch is not using.
struct SppmBuffers {
  DevBuf pix_mem, grid_mem, pair_mem, hist_mem, scan_mem, job, spill, spill_cam;
  DevBuf hits;       // photon hits of a batch of iterations, waiting for their deposits
  DevBuf pq;         // photon queue lengths: 3 rotating sets of one counter per queue segment
  DevBuf cam_terms;  // a camera batch's recorded radiance terms
  WaveBuffers camw;
  std::vector<Pix> px;
  float *ld_snap = nullptr;  // nsets x 3 planes (r, g, b) x npix
  int nsets = 0, cam_batch = 1;
  Grid g{};
  int npix = 0, hash_size = 0;
  size_t pair_cap = 0, hist_cap = 0, scan_cap = 0;
  hipStream_t cam_stream = nullptr;
  hipEvent_t cam_done[2] = {nullptr, nullptr}, iter_done = nullptr;
  std::vector<hipEvent_t> iter_end;  // end of iteration i: iter_end[i % nsets]
  // Photon traversal grids: each bounce's queue length is recorded on the
  // device (striped maxima), copied to pinned memory after the batch, and
  // sizes that bounce's grid in the next batch of the same photon count
  // (waves stride over the queue's chunks, so any grid traces every photon).
  DevBuf segmax;
  unsigned long long *h_segmax = nullptr;
  int *h_npairs = nullptr;  // pinned: the iteration's pair count, copied while emission is queued
  hipEvent_t segmax_copied = nullptr;
  int copied_P = 0;          // photons of the batch whose counts are in flight to h_segmax (0: none)
  int plan_P = 0;            // photons of the batch the plan comes from
  std::vector<int> plan;     // per bounce: photons queued
  // the same for the camera passes' paired traversal, one set per camera
  // batch in flight (two alternate); per bounce: closest-hit and any-hit
  // queue lengths
  DevBuf segmax_cam[2];
  unsigned long long *h_segmax_cam[2] = {nullptr, nullptr};
  hipEvent_t cam_copied[2] = {nullptr, nullptr};
  int cam_copied_N[2] = {0, 0};
  int cam_plan_N = 0;
  std::vector<int> cam_plan;
  ~SppmBuffers() {
    if (cam_stream) { (void)hipStreamSynchronize(cam_stream); (void)hipStreamDestroy(cam_stream); }
    if (segmax_copied) { (void)hipEventSynchronize(segmax_copied); (void)hipEventDestroy(segmax_copied); }
    if (h_segmax) (void)hipHostFree(h_segmax);
    if (h_npairs) (void)hipHostFree(h_npairs);
    for (int k = 0; k < 2; ++k) {
      if (cam_copied[k]) { (void)hipEventSynchronize(cam_copied[k]); (void)hipEventDestroy(cam_copied[k]); }
      if (h_segmax_cam[k]) (void)hipHostFree(h_segmax_cam[k]);
    }
    for (hipEvent_t e : cam_done)
      if (e) (void)hipEventDestroy(e);
    if (iter_done) (void)hipEventDestroy(iter_done);
    for (hipEvent_t e : iter_end)
      if (e) (void)hipEventDestroy(e);
  }
};
// camera paths per batch: up to 2^20 (iterations of 2^18-pixel films at 4 per
// batch), at most 8 iterations
constexpr int kCamBatchPaths = 1 << 20, kCamBatchMax = 8;

void destroy_sppm(SppmBuffers *b) { delete b; }

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ int f2key(float f) {  // order-preserving float -> int
  int i = __float_as_int(f);
  return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float key2f(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7fffffff)); }

// ToGrid (sppm.cpp:40-50): truncation toward zero, then clamp
__device__ __forceinline__ bool to_grid(V3 p, const float *lo, const float *hi, const int *res, int *c) {
  bool in = true;
  float o[3] = {p.x - lo[0], p.y - lo[1], p.z - lo[2]};
  for (int a = 0; a < 3; ++a)
    if (hi[a] > lo[a]) o[a] /= hi[a] - lo[a];
  for (int a = 0; a < 3; ++a) {
    c[a] = (int)(res[a] * o[a]);
    in &= (c[a] >= 0 && c[a] < res[a]);
    c[a] = c[a] < 0 ? 0 : (c[a] > res[a] - 1 ? res[a] - 1 : c[a]);
  }
  return in;
}
// hash (sppm.cpp:52-56); int products wrap, so they are formed in uint32
__device__ __forceinline__ uint32_t grid_hash(int x, int y, int z, int hs) {
  uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
  return h % (uint32_t)hs;
}

__device__ __forceinline__ V3 ld3(const float *a, int i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
__device__ __forceinline__ void st3(float *a, int i, V3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }

// ---------------------------------------------------------------- camera pass
// The camera passes of a batch of nb iterations [c0, c0 + nb) run as one
// wavefront of nb x npix paths: path p is pixel q = p % npix of iteration
// c0 + p / npix.  Iteration ci writes visible-point set ci % nsets.  A
// path's radiance terms (emission, direct light, envmap on escape) are
// recorded in order instead of being added to the pixel's Ld at once: Ld
// accumulates across iterations (sppm.cpp:93-99 adds each term to the
// running sum), so k_cam_replay adds them afterwards, iteration by
// iteration, term by term, in the reference's order.
struct CamTerms {
  float *t;  // 3 planes (r, g, b) of T x N floats: term k of path p at k * N + p
  int *n;    // terms recorded per path
  int T, N;
};

__device__ __forceinline__ int vp_index(int p, int npix, int c0, int nsets) {
  const int j = p / npix;
  return ((c0 + j) % nsets) * npix + (p - j * npix);
}

__global__ void k_cam_raygen(DScene sc, Wave w, CamTerms ct, int c0) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    w.ctrl[C_PATH0] = w.N;
    w.ctrl[C_RAY] = w.N;
    w.ctrl[C_SHADOW] = 0;
    w.ctrl[C_WORK_RAY] = 0;
    w.ctrl[C_WORK_SHADOW] = 0;
  }
  if (p >= w.N) return;
  const int jb = p / w.npix, q = p - jb * w.npix;
  int j = q / w.W, i = q - j * w.W;
  uint32_t index = (uint32_t)(c0 + jb) * (uint32_t)w.npix + (uint32_t)q;  // sppm.cpp:86
  V3 o, d;
  camera_ray(sc, i, j, index, &o, &d);
  put_ray(w, p, o, d, kFltMax);
  w.st[p] = make_float4(1.f, 1.f, 1.f, path_bits(F_ALIVE, 0, 7));
  w.qpath[0][p] = p;
  w.qray[p] = p;
  ct.n[p] = 0;
}

// One camera-pass step (sppm.cpp:93-138) of a batch; px is visible-point set 0.
__global__ __launch_bounds__(256) void k_cam_shade(DScene sc, Wave w, Pix px, CamTerms ct, int cur, int maxDepth,
                                                   int c0, int nsets, unsigned long long *stats) {
  __shared__ int lds[4 * 5];
  const int nxt = cur ^ 1;
  const int count = w.ctrl[C_PATH0 + cur];
  const int N = w.N;
  unsigned long long lost = 0;
  for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    int i = base + threadIdx.x;
    bool act = i < count;
    int p = act ? w.qpath[cur][i] : 0;
    bool want_ray0 = false, want_ray1 = false, want_shadow = false, keep = false;
    if (act) {
      const float4 S = w.st[p];
      int fl = bits_flags(S), depth = bits_bounce(S), dim = bits_dim(S);
      Sp beta = sp3(S.x, S.y, S.z);
      int nt = ct.n[p];
      auto put = [&](Sp x) {
        if (nt < ct.T) {
          const size_t k = (size_t)nt * ct.N + p;
          ct.t[k] = x.r;
          ct.t[(size_t)ct.T * ct.N + k] = x.g;
          ct.t[2 * (size_t)ct.T * ct.N + k] = x.b;
        } else {
          ++lost;
        }
        ++nt;
      };
      if (fl & F_PEND) {
        put(direct_term(w, p, fl));
        fl &= ~(F_PEND | F_SHADOW | F_MIS);
      }
      if (fl & F_ALIVE) {
        fl &= ~F_ALIVE;
        const float4 H = w.hit[p];
        const int hp = __float_as_int(H.w);
        const uint32_t jb = (uint32_t)(p / w.npix);
        const uint32_t index = ((uint32_t)c0 + jb) * (uint32_t)w.npix + (uint32_t)(p - (int)jb * w.npix);
        const float4 D = w.rd[p];
        const V3 d = mk(D.x, D.y, D.z);
        if (hp >= 0) {
          Isect is;
          make_isect(sc, hp, H.y, H.z, &is);
          Bsdf bs;
          make_bsdf(sc.mats + ORION_MAT_FLOATS * (size_t)is.mat, is.ng, is.ns, is.u, is.v, false, &bs);
          V3 wo = neg(d);
          if ((depth == 0 || (fl & F_SPEC)) && is.light >= 0) put(beta * area_emission(sc, is, wo));
          int ref_rays = 0;
          int nf = setup_direct(sc, w, p, bs, is, d, beta, index, &dim, true, &ref_rays);
          fl |= nf;
          want_shadow = (nf & F_SHADOW) != 0;
          want_ray1 = (nf & F_MIS) != 0;
          bool isDiffuse = bsdf_num(bs, DIFF | REFL | TRANS) > 0;
          bool isGlossy = bsdf_num(bs, GLOSSY | REFL | TRANS) > 0;
          if (isDiffuse || (isGlossy && depth == maxDepth - 1)) {
            const int v = vp_index(p, w.npix, c0, nsets);
            px.vpx[v] = is.p.x; px.vpy[v] = is.p.y; px.vpz[v] = is.p.z;
            px.vwox[v] = wo.x; px.vwoy[v] = wo.y; px.vwoz[v] = wo.z;
            st3(px.vng, v, is.ng); st3(px.vns, v, is.ns); st3(px.vu, v, is.u); st3(px.vv, v, is.v);
            px.vmat[v] = is.mat;
            px.vbr[v] = beta.r; px.vbg[v] = beta.g; px.vbb[v] = beta.b;
          } else if (depth < maxDepth - 1) {
            float u1 = sobol(sc.sobol, index, dim++), u0 = sobol(sc.sobol, index, dim++);
            V3 wi = mk(0, 0, 0);
            float pdf = 0;
            int type = 0;
            Sp f = bsdf_sample(bs, &wi, wo, u0, u1, &pdf, ALL, &type);
            if (!(pdf == 0.f || black(f))) {
              fl = (type & SPEC) ? (fl | F_SPEC) : (fl & ~F_SPEC);
              mulby(beta, f * absdot(wi, is.ns) / pdf);
              bool alive = true;
              if (intensity(beta) < 0.25f) {
                float cp = mn(1.0f, intensity(beta));
                if (sobol(sc.sobol, index, dim++) > cp) alive = false;
                else beta = beta / cp;
              }
              if (alive) {
                V3 o = offset_ray_origin(is.p, is.ng, wi);
                put_ray(w, p, o, wi, kFltMax);
                ++depth;
                fl |= F_ALIVE;
                want_ray0 = true;
              }
            }
          }
        } else {  // escaped: every envmap's Le at any depth (sppm.cpp:95-99)
          for (int e = 0; e < sc.n_env; ++e) {
            const float *LR = sc.lights + 32 * (size_t)gld(sc.env_meta + 5 * e);
            put(beta * env_le(sc, LR, d));
          }
        }
      }
      ct.n[p] = nt;
      w.st[p] = make_float4(beta.r, beta.g, beta.b, path_bits(fl, depth, dim));
      keep = (fl & (F_ALIVE | F_PEND)) != 0;
    }
    int a[4];
    block_append4(&w.ctrl[C_PATH0 + nxt], &w.ctrl[C_RAY], &w.ctrl[C_SHADOW], &w.ctrl[C_SHADOW], keep, want_ray0,
                  want_shadow, want_ray1, lds, a);
    if (keep) w.qpath[nxt][a[0]] = p;
    if (want_ray0) w.qray[a[1]] = p;
    if (want_shadow) w.qshadow[a[2]] = p;
    if (want_ray1) w.qshadow[a[3]] = N + p;
  }
  if (lost) stat_add(stats, S_OVERFLOW, lost);
}

// Adds a batch's recorded terms to each pixel's Ld in the reference's order
// (iteration by iteration, term by term) and keeps Ld after each iteration
// (the snapshot its film reads) in set (c0 + j) % nsets of snap (3 planes of
// npix floats per set).
__global__ void k_cam_replay(Pix px, CamTerms ct, float *snap, int npix, int nb, int c0, int nsets) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npix) return;
  Sp Ld = sp3(px.Ldr[q], px.Ldg[q], px.Ldb[q]);
  for (int j = 0; j < nb; ++j) {
    const int p = j * npix + q;
    const int n = min(ct.n[p], ct.T);
    for (int k = 0; k < n; ++k) {
      const size_t o = (size_t)k * ct.N + p;
      acc(Ld, sp3(ct.t[o], ct.t[(size_t)ct.T * ct.N + o], ct.t[2 * (size_t)ct.T * ct.N + o]));
    }
    float *sn = snap + (size_t)((c0 + j) % nsets) * 3 * npix;
    sn[q] = Ld.r; sn[npix + q] = Ld.g; sn[2 * npix + q] = Ld.b;
  }
  px.Ldr[q] = Ld.r; px.Ldg[q] = Ld.g; px.Ldb[q] = Ld.b;
}

// ---------------------------------------------------------------- grid
__global__ void k_grid_init(Grid g) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int big = 0x7f7fffff;  // FLT_MAX as a sortable key
    for (int a = 0; a < 3; ++a) { g.ictl[4 + a] = big; g.ictl[7 + a] = f2key(-3.40282347e+38f); }
    g.ictl[3] = 0;
    g.ictl[10] = 0;
    g.ictl[11] = 0;
  }
}

// bounds = union of expand(vp.p, r) over visible points (sppm.cpp:149-157)
constexpr int kBoundsBlocks = 256;
__global__ __launch_bounds__(256) void k_grid_bounds(Pix px, Grid g, int npix) {
  __shared__ int red[4][8];  // per wave of the 256-thread block
  int lo[3], hi[3], mr = 0, any = 0;
  for (int a = 0; a < 3; ++a) { lo[a] = 0x7f7fffff; hi[a] = f2key(-3.40282347e+38f); }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
    if (black(sp3(px.vbr[p], px.vbg[p], px.vbb[p]))) continue;
    float r = px.radius[p];
    float q[3] = {px.vpx[p], px.vpy[p], px.vpz[p]};
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], f2key(q[a] - r));
      hi[a] = max(hi[a], f2key(q[a] + r));
    }
    mr = max(mr, f2key(r));
    any = 1;
  }
  int vals[8] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], mr, any};
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 8; ++k) {
      int o = __shfl_down(vals[k], off);
      vals[k] = k < 3 ? min(vals[k], o) : max(vals[k], o);
    }
  // block reduction through LDS, then the block's partial to g.part, which
  // k_grid_res reduces: device atomics on these 8 words serialise (one set
  // per block of a 256-block grid took ~18 us, one per wave ~0.3 ms)
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 0)
    for (int k = 0; k < 8; ++k) red[wv][k] = vals[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int v = 1; v < (int)(blockDim.x >> 6); ++v)
    for (int k = 0; k < 8; ++k) vals[k] = k < 3 ? min(vals[k], red[v][k]) : max(vals[k], red[v][k]);
  for (int k = 0; k < 8; ++k) g.part[blockIdx.x * 8 + k] = vals[k];
}

// grid resolution (sppm.cpp:159-163), after the bounds of the nb partials
// of k_grid_bounds; one wave64
__global__ void k_grid_res(Grid g, int nb) {
  const int lane = threadIdx.x;
  int v[8] = {0x7f7fffff, 0x7f7fffff, 0x7f7fffff, f2key(-3.40282347e+38f), f2key(-3.40282347e+38f),
              f2key(-3.40282347e+38f), 0, 0};
  for (int b = lane; b < nb; b += 64)
    for (int k = 0; k < 8; ++k) {
      const int o = g.part[b * 8 + k];
      v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 8; ++k) {
      const int o = __shfl_down(v[k], off);
      v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
    }
  if (lane != 0) return;
  for (int a = 0; a < 3; ++a) { g.ictl[4 + a] = v[a]; g.ictl[7 + a] = v[3 + a]; }
  g.ictl[10] = v[6];
  g.ictl[3] = v[7];
  if (!g.ictl[3]) { g.ictl[0] = g.ictl[1] = g.ictl[2] = 1; return; }
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) { lo[a] = key2f(g.ictl[4 + a]); hi[a] = key2f(g.ictl[7 + a]); }
  float maxRadius = key2f(g.ictl[10]);
  float diag[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  float maxDiag = mx(diag[0], mx(diag[1], diag[2]));
  int baseGridRes = (int)(maxDiag / maxRadius);
  for (int a = 0; a < 3; ++a) {
    int r = (int)(baseGridRes * diag[a] / maxDiag);
    g.ictl[a] = r > 1 ? r : 1;
    g.fctl[a] = lo[a];
    g.fctl[3 + a] = hi[a];
  }
  g.fctl[6] = maxRadius;
}

__device__ __forceinline__ void vp_cells(const Pix &px, const Grid &g, int p, int *pmin, int *pmax) {
  float r = px.radius[p];
  V3 q = mk(px.vpx[p], px.vpy[p], px.vpz[p]);
  to_grid(sub(q, mk(r, r, r)), g.fctl, g.fctl + 3, g.ictl, pmin);
  to_grid(add(q, mk(r, r, r)), g.fctl, g.fctl + 3, g.ictl, pmax);
}

__global__ void k_grid_count(Pix px, Grid g, int npix) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > npix) return;
  int c = 0;
  if (p < npix && g.ictl[3] && !black(sp3(px.vbr[p], px.vbg[p], px.vbb[p]))) {
    int a[3], b[3];
    vp_cells(px, g, p, a, b);
    c = (b[2] - a[2] + 1) * (b[1] - a[1] + 1) * (b[0] - a[0] + 1);
  }
  g.cnt[p] = c;
}

// (key, pixel) pairs in the reference's z, y, x cell order (sppm.cpp:176-193)
// Skips itself when the iteration's pairs (and the one spare entry the
// host keeps) exceed the capacity: the host then grows it and launches the
// emission again.
__global__ void k_grid_emit(Pix px, Grid g, int npix) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix || (int64_t)g.cnt[npix] + 1 > g.pair_cap) return;
  int off = g.cnt[p], end = g.cnt[p + 1];
  if (off == end) return;
  int a[3], b[3];
  vp_cells(px, g, p, a, b);
  for (int z = a[2]; z <= b[2]; ++z)
    for (int y = a[1]; y <= b[1]; ++y)
      for (int x = a[0]; x <= b[0]; ++x) {
        g.keys[0][off] = grid_hash(x, y, z, g.hash_size);
        g.vals[0][off] = p;
        ++off;
      }
}

// ---------------------------------------------------------------- scan (exclusive, int)
constexpr int kScanBlock = 256, kScanItems = 4, kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int block_excl_scan(int v, int *lds, int *total) {
  int lane = lane_id(), wv = threadIdx.x / 64;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[wv] = x;
  __syncthreads();
  int nw = blockDim.x / 64;
  int pre = 0, tot = 0;
  for (int k = 0; k < nw; ++k) {
    int s = lds[k];
    if (k < wv) pre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const int *in, int *out, int *sums, int n) {
  __shared__ int lds[8];
  int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems], s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int tot;
  int pre = block_excl_scan(s, lds, &tot);
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ void k_scan_add(int *out, const int *sums, int n) {
  int i = blockIdx.x * kScanTile + threadIdx.x;
  int add = sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k, i += kScanBlock)
    if (i < n) out[i] += add;
}

// exclusive scan of n ints (in may equal out); tmp must hold scan_tmp_ints(n)
static size_t scan_tmp_ints(size_t n) {
  size_t t = 0;
  while (n > 1) {
    n = (n + kScanTile - 1) / kScanTile;
    t += n + 1;
  }
  return t + 1;
}
static void device_scan(const int *in, int *out, int n, int *tmp, hipStream_t st) {
  int nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles, dim3(std::max(nb, 1)), dim3(kScanBlock), 0, st, in, out, tmp, n);
  if (nb > 1) {
    device_scan(tmp, tmp, nb, tmp + nb + 1, st);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, st, out, tmp, n);
  }
}

// ---------------------------------------------------------------- stable LSD radix sort (8-bit digits)
constexpr int kSortBlock = 256, kSortRows = 16, kSortTile = kSortRows * 64;  // 1024 keys per tile

__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const uint32_t *keys, int n, int shift, int *hist, int ntiles) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kSortTile;
  for (int k = threadIdx.x; k < kSortTile; k += kSortBlock)
    if (base + k < n) atomicAdd(&h[(keys[base + k] >> shift) & 255], 1);
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];  // digit-major
}

// Scatter with stable ranks: row r of the tile (64 keys) computes, per key,
// the lanes holding the same digit with 8 ballots; rows are combined in
// order through an LDS table of per-row digit counts.
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(const uint32_t *kin, const int *vin, uint32_t *kout,
                                                             int *vout, int n, int shift, const int *offs,
                                                             int ntiles) {
  __shared__ int rowcnt[kSortRows][256];
  __shared__ int base_off[256];
  for (int k = threadIdx.x; k < kSortRows * 256; k += kSortBlock) (&rowcnt[0][0])[k] = 0;
  base_off[threadIdx.x] = offs[threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x / 64;
  const int tile0 = blockIdx.x * kSortTile;
  const int rows_per_wave = kSortRows / (kSortBlock / 64);
  int digit[4], rank[4];
  bool valid[4];
  unsigned long long below = lane ? ((~0ull) >> (64 - lane)) : 0ull;
  for (int q = 0; q < rows_per_wave; ++q) {
    int row = wv * rows_per_wave + q;
    int idx = tile0 + row * 64 + lane;
    valid[q] = idx < n;
    int d = valid[q] ? (int)((kin[idx] >> shift) & 255) : 256 + lane;  // invalid lanes never match
    unsigned long long m = __ballot(valid[q]);
    for (int b = 0; b < 8; ++b) {
      unsigned long long bal = __ballot((d >> b) & 1);
      m &= ((d >> b) & 1) ? bal : ~bal;
    }
    digit[q] = d;
    rank[q] = __popcll(m & below);
    if (valid[q] && rank[q] == 0) rowcnt[row][d] = __popcll(m);
  }
  __syncthreads();
  // exclusive prefix over rows, per digit (thread = digit)
  {
    int d = threadIdx.x, acc = 0;
    for (int r = 0; r < kSortRows; ++r) {
      int c = rowcnt[r][d];
      rowcnt[r][d] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int q = 0; q < rows_per_wave; ++q) {
    if (!valid[q]) continue;
    int row = wv * rows_per_wave + q;
    int idx = tile0 + row * 64 + lane;
    int pos = base_off[digit[q]] + rowcnt[row][digit[q]] + rank[q];
    kout[pos] = kin[idx];
    vout[pos] = vin[idx];
  }
}

__global__ void k_cell_clear(Grid g) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.hash_size) {
    g.cell_start[i] = 0; g.cell_end[i] = 0;
    if (g.photon_cell) g.photon_cell[i] = 0;
  }
}
__global__ void k_cell_ranges(Grid g, const uint32_t *keys, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) g.cell_start[k] = i;
  if (i == n - 1 || keys[i + 1] != k) g.cell_end[k] = i + 1;
}
__global__ void k_cell_counts(Grid g, int *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < g.hash_size) out[i] = g.cell_end[i] - g.cell_start[i];
}

// ---------------------------------------------------------------- photon pass
// Emits photon p (sppm.cpp:203-222) into the wave's ray slots; true when it
// carries power.
__device__ __forceinline__ bool photon_emit(DScene sc, Wave w, uint32_t index0, int p) {
  uint32_t index = index0 + (uint32_t)p;  // sppm.cpp:203
  float lightPdf;
  int l = sample_light(sc, sobol(sc.sobol, index, 5), &lightPdf);
  float a1 = sobol(sc.sobol, index, 6), a0 = sobol(sc.sobol, index, 7);
  float b1 = sobol(sc.sobol, index, 8), b0 = sobol(sc.sobol, index, 9);
  V3 o, d, nl;
  float pdfPos, pdfDir;
  Sp Le = light_sample_le(sc, l, a0, a1, b0, b1, &o, &d, &nl, &pdfPos, &pdfDir);
  bool alive = !(pdfPos == 0 || pdfDir == 0 || black(Le));
  Sp beta = sp(0);
  if (alive) {
    beta = Le * absdot(nl, d) / (lightPdf * pdfPos * pdfDir);
    alive = !black(beta);
  }
  put_ray(w, p, o, d, kFltMax);
  w.st[p] = make_float4(beta.r, beta.g, beta.b, path_bits(0, 0, 10));
  return alive;
}

// Photon queues are cut into segments of kPhotonSeg entries, each with its
// own length counter: photon p lives in segment p / kPhotonSeg at every
// bounce, so a block appends to its segment's counter and no counter is
// shared by more than kPhotonSeg / 256 blocks (one counter for the whole
// queue had every block of a launch queue its atomic on one address).
// Set k of the rotating counters is pq + k * nseg.
constexpr int kPhotonSeg = 16384;

// Emits photons [0, P) into queue 0 (counter set 0, zeroed by the host
// beforehand with the other two sets).
__global__ __launch_bounds__(256) void k_photon_init(DScene sc, Wave w, uint32_t index0, int P, int *pq) {
  __shared__ int lds[5];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = (blockIdx.x * blockDim.x) / kPhotonSeg;  // blocks do not straddle segments
  bool alive = false;
  if (p < P) alive = photon_emit(sc, w, index0, p);
  int a = block_append(&pq[sg], alive, lds);
  if (alive) w.qray[(size_t)sg * kPhotonSeg + a] = p;
}

// Photon deposits of one wave (sppm.cpp:230-257), balanced over its lanes.
// Each lane brings one photon and the range [e0, e0 + n) of visible points
// in its cell; the wave walks the concatenation of all 64 ranges, 64 entries
// per step.  Cells near the horizon hold thousands of visible points, so a
// lane-per-photon loop would leave 63 lanes idle behind one such cell.
// ph holds the wave's photons (9 floats each: p, -d, beta), inc the wave's
// inclusive prefix of n.  Returns the deposits this lane made.
__device__ __forceinline__ unsigned long long photon_deposits(const DScene &sc, const Pix &px, const Grid &g,
                                                               int e0, int n, const float *ph, int *ebase,
                                                               int *inc) {
  const int lane = lane_id();
  int x = n;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  const int total = __shfl(x, 63);
  if (total == 0) return 0;
  inc[lane] = x;
  ebase[lane] = e0 - (x - n);  // entry k of the concatenation owned by lane j is ebase[j] + k
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  unsigned long long nd = 0;
  for (int k = lane; k < total; k += 64) {
    int lo = 0, hi = 63;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (inc[mid] > k) hi = mid; else lo = mid + 1;
    }
    const float *f = ph + 9 * lo;
    V3 hp = mk(f[0], f[1], f[2]);
    int q = g.vals[0][ebase[lo] + k];
    float r = px.radius[q];
    V3 vp = mk(px.vpx[q], px.vpy[q], px.vpz[q]);
    if (lensq(sub(vp, hp)) > r * r) continue;
    Bsdf vb;
    make_bsdf(sc.mats + ORION_MAT_FLOATS * (size_t)px.vmat[q], ld3(px.vng, q), ld3(px.vns, q), ld3(px.vu, q),
              ld3(px.vv, q), false, &vb);
    Sp Phi = sp3(f[6], f[7], f[8]) * bsdf_f(vb, mk(f[3], f[4], f[5]), mk(px.vwox[q], px.vwoy[q], px.vwoz[q]), ALL);
    atomicAdd(&px.Phir[q], Phi.r);
    atomicAdd(&px.Phig[q], Phi.g);
    atomicAdd(&px.Phib[q], Phi.b);
    atomicAdd(&px.M[q], 1);
    ++nd;
  }
  __builtin_amdgcn_wave_barrier();  // ph / inc are rewritten by the next call
  return nd;
}

// Photon hits waiting for their iteration's grid.  The reference deposits a
// photon at each bounce after the first as it traces (sppm.cpp:230-257); a
// photon's path depends only on its sample index, not on the visible points,
// so the photons of a batch of iterations are traced together (more rays per
// traversal launch) and every hit is kept until its iteration's grid is
// sorted.  SoA: nine planes (point, -direction, beta) of `total` floats;
// iteration j of the batch owns entries [j * cap, j * cap + cnt[j]).
constexpr int kPhotonBatchRays = 1 << 22;        // photons per batch (up to this many)
constexpr int kPhotonBatchMax = 16;               // iterations per batch
constexpr size_t kPhotonHitBytes = (size_t)2 << 30;  // hit records of a batch
struct PhotonHits {
  float *f;
  int *cnt;
  int cap;      // entries per iteration: photons x (maxDepth - 1), each photon's bounces after the first
  int photons;  // photons per iteration
  size_t total;
};

// One photon bounce (sppm.cpp:225-271) of a batch: the hit is recorded for
// its iteration's deposits (bounces after the first), then the BSDF is
// sampled and the photon continues or ends by Russian roulette.  Bounce d
// reads the queue just traced (qcur, lengths in counter set d % 3) and
// appends its survivors to qnext (set (d + 1) % 3, zeroed before); it
// zeroes set (d + 2) % 3 for the bounce after.  The bps blocks of segment
// sg shade its entries with a stride of bps * 256.  The host alternates the
// two queue buffers, so no launch sits between a bounce and the next trace.
__global__ __launch_bounds__(256) void k_photon_bounce(DScene sc, Wave w, PhotonHits ph, uint32_t index0, int depth,
                                                       const int *qcur, int *qnext, int *pq, int nseg, int bps,
                                                       unsigned long long *stats) {
  __shared__ int lds[5];
  __shared__ int rslot[kPhotonBatchMax];  // per iteration: the block's record count, then its first slot
  const int sg = blockIdx.x / bps, part = blockIdx.x - sg * bps;
  if (sg >= nseg) return;  // the whole block
  const int count = pq[(depth % 3) * nseg + sg];
  if (part == 0 && threadIdx.x == 0) {
    pq[((depth + 2) % 3) * nseg + sg] = 0;
    stat_add(stats, S_CLOSEST, (unsigned long long)count);
    stat_add(stats, S_PHOTON_RAYS, (unsigned long long)count);
  }
  int *const nxt = &pq[((depth + 1) % 3) * nseg + sg];
  const size_t q0 = (size_t)sg * kPhotonSeg;
  const int lane = lane_id();
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  unsigned long long nrec = 0, ovf = 0;
  for (int base = part * blockDim.x; base < count; base += bps * blockDim.x) {
    int i = base + threadIdx.x;
    bool keep = false, rec = false;
    int p = 0, jb = 0;
    V3 hpos = mk(0, 0, 0), d = mk(0, 0, 0);
    Sp beta = sp(0);
    if (i < count) {
      p = qcur[q0 + i];
      const float4 H = w.hit[p];
      const int hp = __float_as_int(H.w);
      if (hp >= 0) {
        Isect is;
        make_isect(sc, hp, H.y, H.z, &is);
        const float4 D = w.rd[p], S = w.st[p];
        d = mk(D.x, D.y, D.z);
        beta = sp3(S.x, S.y, S.z);
        hpos = is.p;
        rec = depth > 0;
        jb = p / ph.photons;
        Bsdf bs;
        make_bsdf(sc.mats + ORION_MAT_FLOATS * (size_t)is.mat, is.ng, is.ns, is.u, is.v, true, &bs);
        uint32_t index = index0 + (uint32_t)p;
        int dim = bits_dim(S);
        float u1 = sobol(sc.sobol, index, dim++), u0 = sobol(sc.sobol, index, dim++);
        V3 wi = mk(0, 0, 0);
        float pdf = 0;
        int flags = 0;
        Sp fr = bsdf_sample(bs, &wi, neg(d), u0, u1, &pdf, ALL, &flags);
        if (!(black(fr) || pdf == 0)) {
          Sp bnew = beta * fr * absdot(wi, is.ns) / pdf;
          float q = mx(0.0f, 1 - intensity(bnew) / intensity(beta));
          if (!(sobol(sc.sobol, index, dim++) < q)) {
            Sp bkeep = bnew / (1 - q);
            V3 o = offset_ray_origin(is.p, is.ng, wi);
            put_ray(w, p, o, wi, kFltMax);
            w.st[p] = make_float4(bkeep.r, bkeep.g, bkeep.b, path_bits(0, 0, dim));
            keep = true;
          }
        }
      }
    }
    // record the hits: slots counted per iteration in LDS, then one global
    // atomic per iteration present in the block (the four to sixteen
    // per-iteration counters are the same few addresses for every block, so
    // atomics per wave would queue behind each other in L2)
    if (threadIdx.x < kPhotonBatchMax) rslot[threadIdx.x] = 0;
    __syncthreads();
    int at = 0;
    uint64_t left = __ballot(rec);
    while (left) {
      const int leader = __ffsll((long long)left) - 1;
      const int j = __shfl(jb, leader);
      const uint64_t mj = __ballot(rec && jb == j);
      int woff = 0;
      if (lane == leader) woff = atomicAdd(&rslot[j], __popcll(mj));
      woff = __shfl(woff, leader);
      if (rec && jb == j) at = woff + __popcll(mj & below);
      left &= ~mj;
    }
    __syncthreads();
    if (threadIdx.x < kPhotonBatchMax) {  // jb < iterations in the batch <= kPhotonBatchMax
      const int c = rslot[threadIdx.x];
      if (c) rslot[threadIdx.x] = atomicAdd(&ph.cnt[threadIdx.x], c);
    }
    __syncthreads();
    if (rec) {
      at += rslot[jb];
      if (at < ph.cap) {
        const size_t k = (size_t)jb * ph.cap + at;
        float *f = ph.f + k;
        f[0] = hpos.x; f[ph.total] = hpos.y; f[2 * ph.total] = hpos.z;
        f[3 * ph.total] = -d.x; f[4 * ph.total] = -d.y; f[5 * ph.total] = -d.z;
        f[6 * ph.total] = beta.r; f[7 * ph.total] = beta.g; f[8 * ph.total] = beta.b;
        ++nrec;
      } else {
        ++ovf;
      }
    }
    int a = block_append(nxt, keep, lds);
    if (keep) qnext[q0 + a] = p;
  }
  for (int off = 32; off > 0; off >>= 1) {
    nrec += __shfl_down(nrec, off);
    ovf += __shfl_down(ovf, off);
  }
  if (lane == 0) {
    stat_add(stats, S_PHOTON_HITS, nrec);
    if (ovf) stat_add(stats, S_OVERFLOW, ovf);
  }
}

// The deposits of iteration jb's photon hits into its visible points
// (sppm.cpp:230-257): per hit its grid cell, then the cell's visible points
// walked by photon_deposits.
__global__ __launch_bounds__(256) void k_photon_deposit(DScene sc, Pix px, Grid g, PhotonHits ph, int jb,
                                                        unsigned long long *stats) {
  __shared__ float phot[256 * 9];
  __shared__ int ebase[256], inc[256];
  const int count = min(ph.cnt[jb], ph.cap);
  const int wv = threadIdx.x >> 6;
  unsigned long long ndep = 0, nlook = 0, nscan = 0;
  for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    int e0 = 0, n = 0;
    if (i < count) {
      const float *f = ph.f + (size_t)jb * ph.cap + i;
      const V3 hp = mk(f[0], f[ph.total], f[2 * ph.total]);
      int c[3];
      if (to_grid(hp, g.fctl, g.fctl + 3, g.ictl, c)) {
        uint32_t h = grid_hash(c[0], c[1], c[2], g.hash_size);
        if (g.photon_cell) atomicAdd(&g.photon_cell[h], 1);  // debug output only
        e0 = g.cell_start[h];
        n = g.cell_end[h] - e0;
        ++nlook;
        nscan += (unsigned long long)n;
        float *q = phot + 9 * threadIdx.x;
        for (int k = 0; k < 9; ++k) q[k] = f[k * ph.total];
      }
    }
    ndep += photon_deposits(sc, px, g, e0, n, phot + 9 * 64 * wv, ebase + 64 * wv, inc + 64 * wv);
  }
  for (int off = 32; off > 0; off >>= 1) {
    ndep += __shfl_down(ndep, off);
    nlook += __shfl_down(nlook, off);
    nscan += __shfl_down(nscan, off);
  }
  if (lane_id() == 0) {
    stat_add(stats, S_DEPOSITS, ndep);
    stat_add(stats, S_LOOKUPS, nlook);
    stat_add(stats, S_SCANNED, nscan);
  }
}

// ---------------------------------------------------------------- update + film
__global__ void k_pixel_update(Pix px, int npix, int *M_out) {  // sppm.cpp:286-310
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  int M = px.M[i];
  if (M_out) M_out[i] = M;
  if (M > 0) {
    float gamma = (float)2 / (float)3;
    float N = px.N[i], r = px.radius[i];
    float Nnew = N + gamma * M;
    float Rnew = r * sqrtf(Nnew / (N + M));
    Sp Phi = sp3(px.Phir[i], px.Phig[i], px.Phib[i]);
    Sp tau = sp3(px.taur[i], px.taug[i], px.taub[i]);
    tau = (tau + sp3(px.vbr[i], px.vbg[i], px.vbb[i]) * Phi) * (Rnew * Rnew) / (r * r);
    px.taur[i] = tau.r; px.taug[i] = tau.g; px.taub[i] = tau.b;
    px.N[i] = Nnew;
    px.radius[i] = Rnew;
    px.M[i] = 0;
    px.Phir[i] = 0; px.Phig[i] = 0; px.Phib[i] = 0;
  }
  px.vbr[i] = 0.f; px.vbg[i] = 0.f; px.vbb[i] = 0.f;
}
__global__ void k_sppm_film(Pix px, int npix, float nIterF, float NpF, float *film, int *count) {  // sppm.cpp:313-320
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  float r = px.radius[i];
  Sp L = sp3(px.Ldr[i], px.Ldg[i], px.Ldb[i]) / nIterF;
  acc(L, sp3(px.taur[i], px.taug[i], px.taub[i]) / (NpF * kPi * r * r));
  film[3 * i] += L.r;
  film[3 * i + 1] += L.g;
  film[3 * i + 2] += L.b;
  count[i] += 1;
}
__global__ void k_pix_init(Pix px, int npix, float r0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  px.radius[i] = r0;
  px.N[i] = 0;
  px.Ldr[i] = px.Ldg[i] = px.Ldb[i] = 0;
  px.taur[i] = px.taug[i] = px.taub[i] = 0;
  px.Phir[i] = px.Phig[i] = px.Phib[i] = 0;
  px.M[i] = 0;
  px.vbr[i] = px.vbg[i] = px.vbb[i] = 0;
}

// ---------------------------------------------------------------- host driver
static int ensure_buffers(orion_scene s, int hash_size) {
  if (!s->sppm) s->sppm = new SppmBuffers();
  SppmBuffers &b = *s->sppm;
  int npix = s->W * s->H;
  if (b.npix != npix) {
    const int Bc = std::max(1, std::min(kCamBatchPaths / std::max(npix, 1), kCamBatchMax));
    const int nsets = 2 * Bc;
    // pixel state, nsets visible-point sets (p/wo, frame, mat, beta), nsets Ld snapshots
    size_t per = 12 + (size_t)nsets * (6 + 12 + 1 + 3) + (size_t)nsets * 3;
    int rc = b.pix_mem.ensure(per * (size_t)npix * 4 + 256);
    if (rc) return rc;
    char *m = (char *)b.pix_mem.ptr;
    size_t off = 0;
    auto take = [&](size_t n) { void *q = m + off; off += n * 4; return q; };
    Pix p;
    p.radius = (float *)take(npix); p.N = (float *)take(npix);
    p.Ldr = (float *)take(npix); p.Ldg = (float *)take(npix); p.Ldb = (float *)take(npix);
    p.taur = (float *)take(npix); p.taug = (float *)take(npix); p.taub = (float *)take(npix);
    p.Phir = (float *)take(npix); p.Phig = (float *)take(npix); p.Phib = (float *)take(npix);
    p.M = (int *)take(npix);
    const size_t sn = (size_t)nsets * npix;
    p.vpx = (float *)take(sn); p.vpy = (float *)take(sn); p.vpz = (float *)take(sn);
    p.vwox = (float *)take(sn); p.vwoy = (float *)take(sn); p.vwoz = (float *)take(sn);
    p.vng = (float *)take(3 * sn); p.vns = (float *)take(3 * sn);
    p.vu = (float *)take(3 * sn); p.vv = (float *)take(3 * sn);
    p.vmat = (int *)take(sn);
    p.vbr = (float *)take(sn); p.vbg = (float *)take(sn); p.vbb = (float *)take(sn);
    b.px.assign(nsets, p);
    for (int k = 0; k < nsets; ++k) {
      Pix &q = b.px[k];
      const size_t o = (size_t)k * npix;
      q.vpx += o; q.vpy += o; q.vpz += o; q.vwox += o; q.vwoy += o; q.vwoz += o;
      q.vng += 3 * o; q.vns += 3 * o; q.vu += 3 * o; q.vv += 3 * o;
      q.vmat += o; q.vbr += o; q.vbg += o; q.vbb += o;
    }
    b.ld_snap = (float *)take(3 * sn);
    if ((rc = b.camw.ensure(Bc * npix))) return rc;
    for (hipEvent_t e : b.iter_end) (void)hipEventDestroy(e);
    b.iter_end.assign(nsets, nullptr);
    for (hipEvent_t &e : b.iter_end) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b.nsets = nsets;
    b.cam_batch = Bc;
    b.npix = npix;
  }
  if (!b.cam_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&b.cam_stream, hipStreamNonBlocking));
    for (hipEvent_t &e : b.cam_done) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&b.iter_done, hipEventDisableTiming));
  }
  if (b.hash_size != hash_size || !b.grid_mem.ptr) {
    size_t ints = 64 + (size_t)(npix + 1) + 3 * (size_t)hash_size + 8 * (size_t)kBoundsBlocks;
    int rc = b.grid_mem.ensure(ints * 4 + 256);
    if (rc) return rc;
    int *m = (int *)b.grid_mem.ptr;
    Grid &g = b.g;
    g.fctl = (float *)m;
    g.ictl = m + 16;
    g.cnt = m + 64;
    g.cell_start = g.cnt + npix + 1;
    g.cell_end = g.cell_start + hash_size;
    g.photon_cell = g.cell_end + hash_size;
    g.part = g.photon_cell + hash_size;
    g.hash_size = hash_size;
    b.hash_size = hash_size;
  }
  size_t need_scan = scan_tmp_ints(npix + 1) + 16;
  if (b.scan_cap < need_scan) {
    int rc = b.scan_mem.ensure(need_scan * 4);
    if (rc) return rc;
    b.scan_cap = need_scan;
    b.g.scan_tmp = (int *)b.scan_mem.ptr;
  }
  return ORION_OK;
}

static int ensure_pairs(SppmBuffers &b, size_t n) {
  if (b.pair_cap < n) {
    size_t cap = std::max(n + n / 4, (size_t)1 << 16);
    int rc = b.pair_mem.ensure(cap * 4 * 4);
    if (rc) return rc;
    uint32_t *m = (uint32_t *)b.pair_mem.ptr;
    b.g.keys[0] = m;
    b.g.keys[1] = m + cap;
    b.g.vals[0] = (int *)(m + 2 * cap);
    b.g.vals[1] = (int *)(m + 3 * cap);
    b.pair_cap = cap;
    b.g.pair_cap = (int64_t)cap;
  }
  size_t tiles = (n + kSortTile - 1) / kSortTile + 1;
  size_t hn = 256 * tiles;
  size_t need_scan = scan_tmp_ints(hn) + 16;
  if (b.hist_cap < hn) {
    int rc = b.hist_mem.ensure(hn * 4);
    if (rc) return rc;
    b.hist_cap = hn;
    b.g.hist = (int *)b.hist_mem.ptr;
  }
  if (b.scan_cap < need_scan) {
    int rc = b.scan_mem.ensure(need_scan * 4);
    if (rc) return rc;
    b.scan_cap = need_scan;
    b.g.scan_tmp = (int *)b.scan_mem.ptr;
  }
  return ORION_OK;
}

// HIP events around the photon pass's traversal (kind 0) and shading (kind
// 1) launches on the stream they run on, when the caller asks for timing.
struct PhotonTimer {
  bool on = false;
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  int mark(int k) {
    if (!on) return ORION_OK;
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    ev.push_back(e);
    kind.push_back(k);
    HIP_TRY(hipEventRecord(e, st));
    return ORION_OK;
  }
  // after the stream has finished: per kind, summed milliseconds and launches
  int collect(double ms[2], int64_t n[2]) {
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float t = 0;
      HIP_TRY(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      ms[kind[i]] += t;
      n[kind[i]]++;
    }
    return ORION_OK;
  }
  ~PhotonTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

}  // namespace orion_amd

using namespace orion_amd;

extern "C" int orion_render_sppm(orion_scene s, int n_iters, int photons, float r0, int maxDepth, int hash_size,
                                 const orion_sppm_debug *dbg, orion_render_stats *stats) {
  if (!s) { set_error("null scene"); return ORION_ERR_ARG; }
  if (n_iters <= 0 || maxDepth <= 0 || !(r0 > 0)) { set_error("bad SPPM parameters"); return ORION_ERR_ARG; }
  if (maxDepth > 60) { set_error("maxDepth above 60 exceeds the Sobol dimension budget"); return ORION_ERR_UNSUPPORTED; }
  const int npix = s->W * s->H;
  if (photons < 0) photons = npix;  // sppm.cpp:329-330
  if (hash_size <= 0) hash_size = npix;  // sppm.cpp:147
  if (hash_size > (1 << 24)) { set_error("hash_size above 2^24"); return ORION_ERR_UNSUPPORTED; }
  // the reference forms photon indices in int: nIterations * nPixels + iter * P + p
  if ((int64_t)n_iters * npix + (int64_t)n_iters * photons >= ((int64_t)1 << 31)) {
    set_error("SPPM sample indices exceed the reference's int range");
    return ORION_ERR_UNSUPPORTED;
  }
  if (s->d.n_lights == 0) { set_error("SPPM needs at least one light"); return ORION_ERR_ARG; }
  orion_ctx c = s->ctx;
  hipStream_t st = c->stream;
  HIP_TRY(hipSetDevice(c->device));
  // a call that returned an error mid-way may have left camera-pass work on
  // the camera stream: let it finish before any buffer is resized
  if (s->sppm && s->sppm->cam_stream) HIP_TRY(hipStreamSynchronize(s->sppm->cam_stream));
  int rc = ensure_buffers(s, hash_size);
  if (rc) return rc;
  SppmBuffers &b = *s->sppm;
  Grid &g = b.g;
  // Photon batches: the photons of B iterations are traced together (on the
  // caustics scene 2^20-ray traversal launches ran at 0.81 ns per ray, 2^22
  // at 0.46, 2^23 at 0.42); their hits wait in b.hits until each
  // iteration's grid is sorted.
  int B = 1;
  PhotonHits ph{};
  if (photons > 0) {
    const int64_t cap = (int64_t)photons * std::max(1, maxDepth - 1);
    if (cap >= ((int64_t)1 << 31)) {
      set_error("SPPM photons x (maxDepth - 1) exceeds 2^31 photon hits per iteration");
      return ORION_ERR_UNSUPPORTED;
    }
    B = std::max(1, std::min({n_iters, kPhotonBatchRays / photons, kPhotonBatchMax}));
    while (B > 1 && (size_t)B * cap * 36 > kPhotonHitBytes) --B;
    if ((rc = b.hits.ensure((size_t)B * cap * 36 + (size_t)B * 4 + 256))) return rc;
    ph.total = (size_t)B * cap;
    ph.f = (float *)b.hits.ptr;
    ph.cnt = (int *)(ph.f + 9 * ph.total);
    ph.cap = (int)cap;
    ph.photons = photons;
  }
  // photon queue segments of the largest batch (queue entries of segment k
  // at k * kPhotonSeg, so the queues hold nseg whole segments)
  const int pseg_max = std::max(1, (B * photons + kPhotonSeg - 1) / kPhotonSeg);
  if ((rc = b.pq.ensure((size_t)3 * pseg_max * 4))) return rc;
  if ((rc = s->wave.ensure(std::max(npix, pseg_max * kPhotonSeg)))) return rc;
  Wave w = s->wave.wave;
  w.ctrl = (int *)s->ctrl.ptr;
  w.stats = (unsigned long long *)s->stats.ptr;
  w.npix = npix;
  w.W = s->W;
  w.H = s->H;
  w.k0 = 0;
  unsigned long long *dstats = (unsigned long long *)s->stats.ptr;
  // one-entry job table for the shared traversal kernels
  Job job1;
  std::memset(&job1, 0, sizeof(job1));
  job1.sc = s->d;
  if ((rc = b.job.ensure(sizeof(Job)))) return rc;
  HIP_TRY(hipMemcpy(b.job.ptr, &job1, sizeof(Job), hipMemcpyHostToDevice));
  const Job *dj = (const Job *)b.job.ptr;
  const int ring = kDefaultRing;
  HIP_TRY(hipMemsetAsync(s->stats.ptr, 0, kStatBytes, st));
  HIP_TRY(hipEventRecord(c->ev0, st));
  const int TB = 256;
  for (const Pix &q : b.px) hipLaunchKernelGGL(k_pix_init, dim3((npix + TB - 1) / TB), dim3(TB), 0, st, q, npix, r0);
  // SPPM's launches trace at most 2^22 rays: 128-entry chunks (two rays per
  // lane) measured 3% faster per iteration than 64, as fast as 256 and 6%
  // faster than 512
  const int chunk = 128;
  if ((rc = b.spill.ensure(trace_spill_ints(s->d.stack_bound, ring, trace_grid(pseg_max, kPhotonSeg, false, chunk)) *
                               4 + 4)))
    return rc;
  // the camera pass traces both its queues in one launch
  if ((rc = b.spill_cam.ensure(
           trace_spill_ints(s->d.stack_bound, ring, trace_pair_grid(1, b.cam_batch * npix, chunk)) * 4 + 4)))
    return rc;
  w.spill = (int *)b.spill.ptr;
  w.spill_cap = std::max(0, s->d.stack_bound - ring);
  const float nIterF = (float)n_iters;
  const float NpF = (float)((uint64_t)n_iters * (uint64_t)photons);
  const bool want_dbg = dbg && dbg->debug_iter >= 0 && dbg->debug_iter < n_iters;
  PhotonTimer ptimer;
  ptimer.on = dbg && dbg->profile;
  ptimer.st = st;
  // camera pass of iteration `ci` (sppm.cpp:84-140) on the camera stream: the
  // wave's radiance arrays are the pixels' Ld; it waits for iteration ci - 2
  // to have finished with visible-point set ci % 2 and Ld snapshot ci % 2
  Wave cw = b.camw.wave;
  cw.ctrl = w.ctrl;
  cw.stats = w.stats;
  cw.npix = npix;
  cw.W = s->W;
  cw.H = s->H;
  cw.k0 = 0;
  cw.spill = (int *)b.spill_cam.ptr;
  cw.spill_cap = w.spill_cap;
  cw.nseg = 1;  // one queue segment; its lengths live in the control block
  cw.cnt_ray = w.ctrl + C_RAY;
  cw.cnt_shadow = w.ctrl + C_SHADOW;
  hipStream_t cst = b.cam_stream;
  const int Bc = std::min(b.cam_batch, n_iters), nsets = b.nsets;
  // a camera path's radiance terms: per bounce the pending direct light and
  // either the hit's emission or every envmap's Le on escape
  CamTerms ct;
  ct.N = b.cam_batch * npix;
  ct.T = (maxDepth + 1) * (1 + std::max(1, s->d.n_env));
  if ((rc = b.cam_terms.ensure((size_t)ct.N * (3 * (size_t)ct.T + 1) * 4))) return rc;
  ct.t = (float *)b.cam_terms.ptr;
  ct.n = (int *)(ct.t + 3 * (size_t)ct.T * ct.N);
  HIP_TRY(hipEventRecord(b.iter_done, st));
  // camera passes of iterations [c0, c0 + nb): they overwrite the sets of
  // iterations c0 - nsets .. c0 + nb - 1 - nsets, so wait for the last of those
  for (int k = 0; k < 2; ++k) {
    if ((rc = b.segmax_cam[k].ensure(kSegMaxBytes))) return rc;
    if (!b.h_segmax_cam[k]) HIP_TRY(hipHostMalloc((void **)&b.h_segmax_cam[k], kSegMaxBytes, hipHostMallocDefault));
    if (!b.cam_copied[k]) HIP_TRY(hipEventCreateWithFlags(&b.cam_copied[k], hipEventDisableTiming));
  }
  const int ncam = std::min(maxDepth + 1, kSegMaxBounces);  // camera bounces whose queue lengths are recorded
  auto camera_batch = [&](int c0, int nb) -> int {
    const int e = c0 + nb - 1 - nsets;
    HIP_TRY(hipStreamWaitEvent(cst, e >= 0 ? b.iter_end[e % nsets] : b.iter_done, 0));
    Wave cb = cw;
    cb.N = nb * npix;
    cb.seg = nb * npix;
    // traversal grids sized from the queue lengths of the camera batch two
    // back (the one that used this set of counters; the main stream has
    // waited for it before the host got here, so its copy is complete)
    const int slot = (c0 / Bc) & 1;
    if (b.cam_copied_N[slot] > 0) {
      HIP_TRY(hipEventSynchronize(b.cam_copied[slot]));
      b.cam_plan.assign(2 * ncam, 0);
      for (int q = 0; q < 2 * ncam; ++q)
        for (int k = 0; k < kStatStripes; ++k)
          b.cam_plan[q] = std::max(b.cam_plan[q], (int)std::min<unsigned long long>(
                                                      b.h_segmax_cam[slot][(size_t)q * kStatStripes + k], 1u << 30));
      b.cam_plan_N = b.cam_copied_N[slot];
      b.cam_copied_N[slot] = 0;
    }
    auto cam_cps = [&](int q) -> int {  // chunk positions for queue q: its length + a quarter + 256
      if (b.cam_plan_N != cb.N || q >= (int)b.cam_plan.size()) return 0;
      const int64_t need = (int64_t)b.cam_plan[q] + b.cam_plan[q] / 4 + 256;
      return (int)std::max<int64_t>(1, (need + chunk - 1) / chunk);
    };
    cb.segmax = (unsigned long long *)b.segmax_cam[slot].ptr;
    HIP_TRY(hipMemsetAsync(b.segmax_cam[slot].ptr, 0, (size_t)2 * ncam * kStatStripes * 8, cst));
    hipLaunchKernelGGL(k_cam_raygen, dim3((cb.N + TB - 1) / TB), dim3(TB), 0, cst, s->d, cb, ct, c0);
    int cur = 0;
    const int blocks = std::min((cb.N + 255) / 256, 8192);
    for (int it = 0; it <= maxDepth; ++it) {
      // closest-hit + any-hit queues in one launch; small planned queues with
      // one ray per lane (64-entry chunks, grids within the full ones)
      int ch = chunk, cc = cam_cps(2 * it), ca = cam_cps(2 * it + 1);
      if (cc > 0 && ca > 0 && (int64_t)cc * chunk <= cb.seg / 2 && (int64_t)ca * chunk <= cb.seg) {
        ch = 64;
        cc = (int)((((int64_t)b.cam_plan[2 * it] + b.cam_plan[2 * it] / 4 + 256) + ch - 1) / ch);
        ca = (int)((((int64_t)b.cam_plan[2 * it + 1] + b.cam_plan[2 * it + 1] / 4 + 256) + ch - 1) / ch);
      }
      launch_trace_pair(dj, cb, ring, ch, false, cst, cc, ca, it < ncam ? it : -1);
      launch_between(cb, cur ^ 1, cst);
      // the shading grid: the bounce's live paths are at most its closest-hit
      // plus any-hit queue lengths (a grid-stride loop shades any count)
      int cblocks = blocks;
      if (b.cam_plan_N == cb.N && 2 * it + 1 < (int)b.cam_plan.size()) {
        const int64_t np = (int64_t)b.cam_plan[2 * it] + b.cam_plan[2 * it + 1];
        cblocks = (int)std::min<int64_t>(blocks, (np + np / 4 + 256 + 255) / 256);
      }
      hipLaunchKernelGGL(k_cam_shade, dim3(cblocks), dim3(256), 0, cst, s->d, cb, b.px[0], ct, cur, maxDepth, c0,
                         nsets, dstats);
      cur ^= 1;
    }
    hipLaunchKernelGGL(k_cam_replay, dim3((npix + TB - 1) / TB), dim3(TB), 0, cst, b.px[0], ct, b.ld_snap, npix, nb,
                       c0, nsets);
    HIP_TRY(hipMemcpyAsync(b.h_segmax_cam[slot], b.segmax_cam[slot].ptr, (size_t)2 * ncam * kStatStripes * 8,
                           hipMemcpyDeviceToHost, cst));
    HIP_TRY(hipEventRecord(b.cam_copied[slot], cst));
    b.cam_copied_N[slot] = cb.N;
    HIP_TRY(hipEventRecord(b.cam_done[(c0 / Bc) & 1], cst));
    return ORION_OK;
  };
  // photon paths of iterations [i0, i0 + nb) (sppm.cpp:199-273 without the
  // deposits): photon p of the batch is photon p % photons of iteration
  // i0 + p / photons, sample index nIterations * nPixels + i0 * photons + p
  // (sppm.cpp:203)
  if ((rc = b.segmax.ensure(kSegMaxBytes))) return rc;
  if (!b.h_segmax) HIP_TRY(hipHostMalloc((void **)&b.h_segmax, kSegMaxBytes, hipHostMallocDefault));
  if (!b.h_npairs) HIP_TRY(hipHostMalloc((void **)&b.h_npairs, sizeof(int), hipHostMallocDefault));
  if (!b.segmax_copied) HIP_TRY(hipEventCreateWithFlags(&b.segmax_copied, hipEventDisableTiming));
  const int nrec = std::min(maxDepth, kSegMaxBounces);
  auto photon_batch = [&](int i0, int nb) -> int {
    const int P = nb * photons;
    const int nseg = (P + kPhotonSeg - 1) / kPhotonSeg;
    int *const pq = (int *)b.pq.ptr;
    Wave pw = w;
    pw.N = P;
    pw.seg = kPhotonSeg;
    pw.nseg = nseg;
    pw.segmax = (unsigned long long *)b.segmax.ptr;
    // the previous batch's per-bounce photon counts (copied a batch ago, so
    // the event has long completed) become this size's plan
    if (b.copied_P > 0) {
      HIP_TRY(hipEventSynchronize(b.segmax_copied));
      b.plan.assign(nrec, 0);
      for (int d = 0; d < nrec; ++d)
        for (int k = 0; k < kStatStripes; ++k)
          b.plan[d] = std::max(b.plan[d], (int)std::min<unsigned long long>(b.h_segmax[(size_t)d * kStatStripes + k], 1u << 30));
      b.plan_P = b.copied_P;
      b.copied_P = 0;
    }
    auto plan_cps = [&](int d) -> int {  // chunk positions per segment for bounce d: its longest segment + a quarter + 256
      if (b.plan_P != P || d >= (int)b.plan.size()) return 0;
      const int64_t need = (int64_t)b.plan[d] + b.plan[d] / 4 + 256;
      return (int)std::max<int64_t>(1, (need + chunk - 1) / chunk);
    };
    HIP_TRY(hipMemsetAsync(b.segmax.ptr, 0, (size_t)nrec * kStatStripes * 8, st));
    const uint32_t index0 = (uint32_t)(n_iters * npix + i0 * photons);
    HIP_TRY(hipMemsetAsync(ph.cnt, 0, (size_t)nb * 4, st));
    HIP_TRY(hipMemsetAsync(pq, 0, (size_t)3 * nseg * 4, st));
    hipLaunchKernelGGL(k_photon_init, dim3((P + TB - 1) / TB), dim3(TB), 0, st, s->d, pw, index0, P, pq);
    int *qbuf[2] = {w.qray, w.qpath[0]};  // bounce d traces qbuf[d & 1]
    const int full_bps = (std::min(kPhotonSeg, P) + 255) / 256;  // shading blocks per segment of a full grid
    for (int depth = 0; depth < maxDepth; ++depth) {
      Wave tw = pw;
      tw.qray = qbuf[depth & 1];
      tw.cnt_ray = pq + (depth % 3) * nseg;
      int rc2;
      if ((rc2 = ptimer.mark(0))) return rc2;
      // a planned bounce whose segments hold at most half a segment traces
      // one photon per lane (64-entry chunks: its grid stays within the
      // 128-entry full grid the spill area holds); its latency, not its
      // work, sets the launch's time
      int ch = chunk, cps = plan_cps(depth);
      if (cps > 0 && (int64_t)cps * chunk <= kPhotonSeg / 2) {
        ch = 64;
        cps = (int)((((int64_t)b.plan[depth] + b.plan[depth] / 4 + 256) + ch - 1) / ch);
      }
      launch_trace_photons(dj, tw, ring, ch, st, cps, depth < nrec ? depth : -1);
      if ((rc2 = ptimer.mark(0)) || (rc2 = ptimer.mark(1))) return rc2;
      // the bounce's shading blocks per segment from the same plan (they
      // stride over their segment, so any count shades every photon)
      const int bps = b.plan_P == P && depth < (int)b.plan.size()
                          ? std::max(1, std::min(full_bps, (b.plan[depth] + b.plan[depth] / 4 + 256 + 255) / 256))
                          : full_bps;
      hipLaunchKernelGGL(k_photon_bounce, dim3(nseg * bps), dim3(256), 0, st, s->d, tw, ph, index0, depth,
                         qbuf[depth & 1], qbuf[(depth + 1) & 1], pq, nseg, bps, dstats);
      if ((rc2 = ptimer.mark(1))) return rc2;
    }
    HIP_TRY(hipMemcpyAsync(b.h_segmax, b.segmax.ptr, (size_t)nrec * kStatStripes * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(b.segmax_copied, st));
    b.copied_P = P;
    return ORION_OK;
  };
  if ((rc = camera_batch(0, Bc))) return rc;
  int batch0 = 0, nbatch = 0;
  for (int iter = 0; iter < n_iters; ++iter) {
    // a new photon batch first: the main stream is idle here (the host
    // waited for the last iteration's pair count), and photons do not
    // depend on any camera pass
    if (photons > 0 && iter >= batch0 + nbatch) {
      batch0 = iter;
      nbatch = std::min(B, n_iters - iter);
      if ((rc = photon_batch(batch0, nbatch))) return rc;
    }
    // the next camera batch overlaps this batch's grids, deposits and photons
    if (iter % Bc == 0 && iter + Bc < n_iters && (rc = camera_batch(iter + Bc, std::min(Bc, n_iters - iter - Bc))))
      return rc;
    HIP_TRY(hipStreamWaitEvent(st, b.cam_done[(iter / Bc) & 1], 0));
    const Pix &px = b.px[iter % nsets];
    // --- grid
    hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(64), 0, st, g);
    const int bblocks = std::max(1, std::min((npix + 255) / 256, kBoundsBlocks));
    hipLaunchKernelGGL(k_grid_bounds, dim3(bblocks), dim3(256), 0, st, px, g, npix);
    hipLaunchKernelGGL(k_grid_res, dim3(1), dim3(64), 0, st, g, bblocks);
    hipLaunchKernelGGL(k_grid_count, dim3((npix + 1 + TB - 1) / TB), dim3(TB), 0, st, px, g, npix);
    device_scan(g.cnt, g.cnt, npix + 1, g.scan_tmp, st);
    HIP_TRY(hipMemcpyAsync(b.h_npairs, g.cnt + npix, 4, hipMemcpyDeviceToHost, st));
    // the pair emission (into the current capacity) and the cell clear need
    // no pair count on the host: queued before the host waits for it
    const size_t cap0 = b.pair_cap;
    hipLaunchKernelGGL(k_grid_emit, dim3((npix + TB - 1) / TB), dim3(TB), 0, st, px, g, npix);
    // per-cell photon counts are kept only for the iteration the debug output
    // asks for: caustic photons land in few cells, and these atomics then
    // serialise
    const bool cell_counts = want_dbg && iter == dbg->debug_iter && dbg->photon_cell_count;
    {
      Grid gc = g;
      if (!cell_counts) gc.photon_cell = nullptr;
      hipLaunchKernelGGL(k_cell_clear, dim3((hash_size + TB - 1) / TB), dim3(TB), 0, st, gc);
    }
    HIP_TRY(hipStreamSynchronize(st));
    const int n_pairs = *b.h_npairs;
    if (n_pairs < 0) { set_error("visible-point cell pairs overflow"); return ORION_ERR_UNSUPPORTED; }
    if ((rc = ensure_pairs(b, (size_t)n_pairs + 1))) return rc;
    if ((size_t)n_pairs + 1 > cap0)  // the queued emission skipped itself
      hipLaunchKernelGGL(k_grid_emit, dim3((npix + TB - 1) / TB), dim3(TB), 0, st, px, g, npix);
    // the deposits' grid: taken after ensure_pairs, which may have moved the
    // pairs and histograms
    Grid gi = g;
    if (!cell_counts) gi.photon_cell = nullptr;
    int bits = 1;
    while ((1 << bits) < hash_size) ++bits;
    int ntiles = (n_pairs + kSortTile - 1) / kSortTile;
    int src = 0;
    if (n_pairs > 0) {
      for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(kSortBlock), 0, st, g.keys[src], n_pairs, shift, g.hist,
                           ntiles);
        device_scan(g.hist, g.hist, 256 * ntiles, g.scan_tmp, st);
        hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(kSortBlock), 0, st, g.keys[src], g.vals[src],
                           g.keys[src ^ 1], g.vals[src ^ 1], n_pairs, shift, g.hist, ntiles);
        src ^= 1;
      }
    }
    if (src == 1) {  // keep the sorted pairs in slot 0
      HIP_TRY(hipMemcpyAsync(g.keys[0], g.keys[1], (size_t)n_pairs * 4, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(g.vals[0], g.vals[1], (size_t)n_pairs * 4, hipMemcpyDeviceToDevice, st));
    }
    if (n_pairs > 0)
      hipLaunchKernelGGL(k_cell_ranges, dim3((n_pairs + TB - 1) / TB), dim3(TB), 0, st, g, g.keys[0], n_pairs);
    if (want_dbg && iter == dbg->debug_iter) {
      if (dbg->vp_cell_count) {
        DevBuf tmp;
        if ((rc = tmp.ensure((size_t)hash_size * 4))) return rc;
        hipLaunchKernelGGL(k_cell_counts, dim3((hash_size + TB - 1) / TB), dim3(TB), 0, st, g, (int *)tmp.ptr);
        HIP_TRY(hipMemcpyAsync(dbg->vp_cell_count, tmp.ptr, (size_t)hash_size * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
      }
      int ictl[16];
      float fctl[16];
      HIP_TRY(hipMemcpyAsync(ictl, g.ictl, sizeof(ictl), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipMemcpyAsync(fctl, g.fctl, sizeof(fctl), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (dbg->grid_info) {
        dbg->grid_info[0] = ictl[0]; dbg->grid_info[1] = ictl[1]; dbg->grid_info[2] = ictl[2];
        dbg->grid_info[3] = n_pairs;
      }
      if (dbg->grid_bounds)
        for (int a = 0; a < 6; ++a) dbg->grid_bounds[a] = fctl[a];
      if (dbg->pairs) {
        if (n_pairs > dbg->max_pairs) {
          set_error("debug pair buffer holds " + std::to_string(dbg->max_pairs) + " pairs, the grid has " +
                    std::to_string(n_pairs));
          return ORION_ERR_ARG;
        }
        std::vector<int> k((size_t)n_pairs), v((size_t)n_pairs);
        if (n_pairs > 0) {
          HIP_TRY(hipMemcpyAsync(k.data(), g.keys[0], (size_t)n_pairs * 4, hipMemcpyDeviceToHost, st));
          HIP_TRY(hipMemcpyAsync(v.data(), g.vals[0], (size_t)n_pairs * 4, hipMemcpyDeviceToHost, st));
          HIP_TRY(hipStreamSynchronize(st));
        }
        for (int i = 0; i < n_pairs; ++i) { dbg->pairs[2 * i] = k[i]; dbg->pairs[2 * i + 1] = v[i]; }
      }
      if (dbg->cell_range) {
        std::vector<int> cs((size_t)hash_size), ce((size_t)hash_size);
        HIP_TRY(hipMemcpyAsync(cs.data(), g.cell_start, (size_t)hash_size * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(ce.data(), g.cell_end, (size_t)hash_size * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (int h = 0; h < hash_size; ++h) { dbg->cell_range[2 * h] = cs[h]; dbg->cell_range[2 * h + 1] = ce[h]; }
      }
    }
    // --- deposits of this iteration's photon hits (sppm.cpp:230-257)
    if (photons > 0) {
      if ((rc = ptimer.mark(1))) return rc;
      const int dblocks = std::max(1, std::min((ph.cap + 255) / 256, 8192));
      hipLaunchKernelGGL(k_photon_deposit, dim3(dblocks), dim3(256), 0, st, s->d, px, gi, ph, iter - batch0, dstats);
      if ((rc = ptimer.mark(1))) return rc;
    }
    if (want_dbg && iter == dbg->debug_iter && dbg->photon_cell_count)
      HIP_TRY(hipMemcpyAsync(dbg->photon_cell_count, g.photon_cell, (size_t)hash_size * 4, hipMemcpyDeviceToHost, st));
    // --- update + film
    DevBuf mdbg;
    int *M_out = nullptr;
    if (want_dbg && iter == dbg->debug_iter && dbg->M) {
      if ((rc = mdbg.ensure((size_t)npix * 4))) return rc;
      M_out = (int *)mdbg.ptr;
    }
    hipLaunchKernelGGL(k_pixel_update, dim3((npix + TB - 1) / TB), dim3(TB), 0, st, px, npix, M_out);
    Pix pf = px;  // Ld as the camera pass left it
    pf.Ldr = b.ld_snap + (size_t)(iter % nsets) * 3 * npix;
    pf.Ldg = pf.Ldr + npix;
    pf.Ldb = pf.Ldg + npix;
    hipLaunchKernelGGL(k_sppm_film, dim3((npix + TB - 1) / TB), dim3(TB), 0, st, pf, npix, nIterF, NpF,
                       (float *)s->film.ptr, (int *)s->count.ptr);
    if (M_out) {
      HIP_TRY(hipMemcpyAsync(dbg->M, M_out, (size_t)npix * 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    HIP_TRY(hipEventRecord(b.iter_end[iter % nsets], st));
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(c->ev1, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  unsigned long long h[S_COUNT];
  if ((rc = read_striped_stats(s->stats.ptr, h))) return rc;
  if ((rc = check_overflow(h[S_OVERFLOW]))) return rc;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->closest_rays = (int64_t)h[S_CLOSEST];
    stats->shadow_rays = (int64_t)h[S_SHADOW];
    stats->ref_equiv_rays = (int64_t)(h[S_CLOSEST] + h[S_SHADOW] + h[S_REFEQ]);
    stats->photons = (int64_t)n_iters * photons;
    stats->paths = (int64_t)n_iters * npix;
    stats->deposits = (int64_t)h[S_DEPOSITS];
    stats->ms = ms;
    stats->node_visits_closest = (int64_t)h[S_NODES_C];
    stats->tri_tests_closest = (int64_t)h[S_TRIS_C];
    stats->node_visits_shadow = (int64_t)h[S_NODES_S];
    stats->tri_tests_shadow = (int64_t)h[S_TRIS_S];
    stats->photon_rays = (int64_t)h[S_PHOTON_RAYS];
    stats->node_visits_photon = (int64_t)h[S_NODES_P];
    stats->tri_tests_photon = (int64_t)h[S_TRIS_P];
    stats->photon_lookups = (int64_t)h[S_LOOKUPS];
    stats->photon_vp_scanned = (int64_t)h[S_SCANNED];
    double pms[2] = {0, 0};
    int64_t pn[2] = {0, 0};
    if ((rc = ptimer.collect(pms, pn))) return rc;
    stats->photon_trace_ms = pms[0];
    stats->photon_trace_launches = pn[0];
    stats->photon_shade_ms = pms[1];
    stats->photon_shade_launches = pn[1];
    stats->photon_hits = (int64_t)h[S_PHOTON_HITS];
    stats->photon_batch = photons > 0 ? B : 0;
  }
  return ORION_OK;
}
