// Device runtime: contexts, scene upload (BVH + shading data + Sobol table
// in HBM), the wavefront render loops and the device half of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"
#include "bvh.h"
#include "host_scene.h"
#include "kernels.h"
#include "runtime.h"

#define ORION_SOBOL_STORAGE static const
#include "sobol_matrices.inc"

namespace {
// Sobol generator matrices as byte tables: entry [dim][j][v] is the XOR of
// the columns 8 j + b of dimension dim for the set bits b of v, so a sample
// is the XOR of four independent lookups (one per byte of the index)
// instead of a chain over the index's set bits.
const std::vector<uint32_t> &sobol_byte_tables() {
  static const std::vector<uint32_t> tab = [] {
    std::vector<uint32_t> t((size_t)1024 * 4 * 256);
    for (int dim = 0; dim < 1024; ++dim)
      for (int j = 0; j < 4; ++j)
        for (int v = 0; v < 256; ++v) {
          uint32_t r = 0;
          for (int b = 0; b < 8; ++b)
            if (v >> b & 1) r ^= orion_sobol_matrices[dim * 32 + 8 * j + b];
          t[((size_t)dim * 4 + j) * 256 + v] = r;
        }
    return t;
  }();
  return tab;
}
}  // namespace

namespace orion_amd {

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                     \
      return ORION_ERR_HIP;                                                             \
    }                                                                                   \
  } while (0)

int dev_alloc(void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();  // a caller may retry smaller: leave no error for the next launch check
    set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? ORION_ERR_OOM : ORION_ERR_HIP;
  }
  return ORION_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}
int DevBuf::ensure(size_t b) {
  if (b <= bytes && ptr) return ORION_OK;
  release();
  int rc = dev_alloc(&ptr, b);
  if (rc == ORION_OK) bytes = b;
  return rc;
}

// ---------------------------------------------------------------- wave buffers
int WaveBuffers::ensure(int N) {
  if (N <= cap && mem.ptr) return ORION_OK;
  // bytes per path: rays 2 x 16 x 3 slots, hit 16, vis 2 x 4, state 16 + 16,
  // pending 3 x 16, queues 4 x (2 + 1 + 2); the float4 arrays first (16-byte
  // aligned), then the int arrays
  const size_t per = 96 + 16 + 8 + 32 + 48 + 20;
  const size_t nseg = ((size_t)N + kSegMin - 1) / kSegMin + 1;
  int rc = mem.ensure(per * (size_t)N + 9 * nseg * 4 + 256);
  if (rc) return rc;
  cap = N;
  char *b = (char *)mem.ptr;
  size_t off = 0;
  auto take = [&](size_t bytes) { void *q = b + off; off += bytes; return q; };
  auto take4 = [&](size_t n) { return (float4 *)take(n * sizeof(float4)); };
  Wave &w = wave;
  w.ro = take4(3 * (size_t)N);
  w.rd = take4(3 * (size_t)N);
  w.hit = take4(N);
  w.st = take4(N);
  w.rad = take4(N);
  w.pa = take4(N);
  w.pm = take4(N);
  w.pn = take4(N);
  auto takei = [&](size_t n) { return (int *)take(n * 4); };
  w.vis = takei(2 * (size_t)N);
  w.N = N;  // slot stride
  w.qpath[0] = takei(N); w.qpath[1] = takei(N);
  w.qray = takei(N);
  w.qshadow = takei(2 * (size_t)N);
  choose_segments(N, &w.seg, &w.nseg);
  w.cnt = takei(9 * nseg);
  w.cnt_stride = (int)nseg;
  w.cnt_ray = w.cnt + nseg;
  w.cnt_shadow = w.cnt + 2 * nseg;
  w.shade_parts = 1;
  w.seg_job = nullptr;
  w.seg_perm = nullptr;
  return ORION_OK;
}

int read_striped_stats(const void *dev, unsigned long long *out) {
  std::vector<unsigned long long> h((size_t)S_COUNT * kStatStripes);
  hipError_t e = hipMemcpy(h.data(), dev, kStatBytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error(std::string("hipMemcpy (statistics): ") + hipGetErrorString(e));
    return ORION_ERR_HIP;
  }
  for (int s = 0; s < S_COUNT; ++s) {
    unsigned long long v = 0;
    for (int k = 0; k < kStatStripes; ++k) v += h[(size_t)s * kStatStripes + k];
    out[s] = v;
  }
  return ORION_OK;
}

// ---------------------------------------------------------------- allocation-free intersection
// Scratch of orion_intersect_closest_noalloc: the traversal spill area of its
// fixed grid (kNoAllocBlocks one-wave blocks x 64 lanes x kNoAllocSpill
// stack entries beyond the LDS ring: scenes whose stack bound is at most
// kNoAllocSpill + kDefaultRing) and, in its own 256-byte line, the count of
// dropped stack entries.  32 MiB + 256 B per GPU, held until the process
// ends.  Host code never uses the array's name: orion_ctx_create stores this
// device's address of it (hipGetSymbolAddress) in orion_ctx::scratch.
constexpr int kNoAllocBlocks = 4096;
constexpr int kNoAllocSpill = 32;
constexpr size_t kNoAllocSpillInts = (size_t)kNoAllocBlocks * kTraceBlock * kNoAllocSpill;
static __device__ __align__(256) int scratch[kNoAllocSpillInts + 64];

__global__ void k_zero_ints(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int check_overflow(unsigned long long dropped) {
  if (!dropped) return ORION_OK;
  set_error("BVH traversal stack overflow: " + std::to_string(dropped) +
            " stack entries dropped (the tree is deeper than its stack bound); results are invalid");
  return ORION_ERR_UNSUPPORTED;
}

}  // namespace orion_amd

using namespace orion_amd;

// ======================================================================= C ABI (device)
extern "C" {

int orion_ctx_create(int device, orion_ctx *out) {
  if (!out) { set_error("null out"); return ORION_ERR_ARG; }
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) { set_error("no such device"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(device));
  auto *c = new orion_ctx_s();
  c->device = device;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->n_cu = prop.multiProcessorCount;
  c->arch = prop.gcnArchName;
  if (c->arch.rfind("gfx950", 0) != 0) {
    delete c;
    set_error("orion_amd is built for gfx950 only; device reports " + std::string(prop.gcnArchName));
    return ORION_ERR_UNSUPPORTED;
  }
  // this device's copy of the allocation-free entry point's scratch array
  hipError_t e = hipGetSymbolAddress((void **)&c->scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !c->scratch) {
    delete c;
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return ORION_ERR_HIP;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e != hipSuccess) {
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    set_error(std::string("context stream / events: ") + hipGetErrorString(e));
    return ORION_ERR_HIP;
  }
  *out = c;
  return ORION_OK;
}

int orion_ctx_destroy(orion_ctx c) {
  if (!c) return ORION_OK;
  if (c->children.load() > 0) {
    set_error("context still has scenes or servers: destroy them first");
    return ORION_ERR_BUSY;
  }
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  orion_amd::destroy_runner(c->runner);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return ORION_OK;
}

int orion_ctx_synchronize(orion_ctx c) {
  if (!c) { set_error("null ctx"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ORION_OK;
}

namespace {
// Every float bit pattern with |x| in [2^-125, 2^125): rcp_rn(x) against
// the IEEE quotient 1 / x; mismatches summed per thread, one atomic each.
__global__ void k_selfcheck_rcp(unsigned long long *mismatches) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long bad = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    const float x = __uint_as_float((uint32_t)i);
    const float ax = fabsf(x);
    if (!(ax >= 0x1p-125f && ax < 0x1p125f)) continue;
    bad += __float_as_uint(orion_amd::dev::rcp_rn(x)) != __float_as_uint(1.0f / x) ? 1 : 0;
  }
  if (bad) atomicAdd(mismatches, bad);
}
}  // namespace

int orion_selfcheck_rcp(orion_ctx c, uint64_t *mismatches) {
  if (!c || !mismatches) { set_error("null argument"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(c->device));
  unsigned long long *d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(unsigned long long)));
  unsigned long long h = 0;
  hipError_t e = hipMemcpyAsync(d, &h, sizeof h, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_selfcheck_rcp, dim3(16384), dim3(256), 0, c->stream, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  HIP_TRY(e);
  *mismatches = h;
  return ORION_OK;
}

int orion_scene_upload(orion_ctx c, orion_hscene hs, orion_scene *out) {
  if (!c || !hs || !out) { set_error("null argument"); return ORION_ERR_ARG; }
  const HostScene &s = *reinterpret_cast<const HostScene *>(hs);
  int W = (int)s.camera[14], H = (int)s.camera[15];
  if (W <= 0 || H <= 0) { set_error("camera record has no film size"); return ORION_ERR_ARG; }
  int n = s.n_tris();
  if (n > (1 << 28) / 8) { set_error("too many triangles for the leaf encoding"); return ORION_ERR_UNSUPPORTED; }
  for (int i = 0; i < n; ++i) {
    if (s.tri_mat[i] < 0 || s.tri_mat[i] >= s.n_mats()) { set_error("triangle material out of range"); return ORION_ERR_ARG; }
    if (s.tri_light[i] >= s.n_lights()) { set_error("triangle light out of range"); return ORION_ERR_ARG; }
  }
  for (int m = 0; m < s.n_mats(); ++m) {
    float t = s.mats[(size_t)kMatFloats * m];
    if (!(t == MAT_MATTE || t == MAT_METAL || t == MAT_GLASS || t == MAT_MIRROR || t == MAT_PLASTIC || t == MAT_UBER)) {
      set_error("material " + std::to_string(m) + " has an unknown type");
      return ORION_ERR_ARG;
    }
  }
  HIP_TRY(hipSetDevice(c->device));
  auto sc = new orion_scene_s();
  sc->ctx = c;
  sc->W = W;
  sc->H = H;
  auto fail = [&](int rc) { delete sc; return rc; };
  auto t0 = std::chrono::steady_clock::now();
  build_bvh(s.tri_pn.data(), n, &sc->bvh);
  auto t1 = std::chrono::steady_clock::now();
  // shading records in primitive order: p0 p1 p2 n0 n1 n2 mat light (80 B)
  std::vector<float> shade(20 * (size_t)std::max(n, 1), 0.f);
  for (int i = 0; i < n; ++i) {
    std::memcpy(&shade[20 * (size_t)i], &s.tri_pn[18 * (size_t)i], 18 * sizeof(float));
    std::memcpy(&shade[20 * (size_t)i + 18], &s.tri_mat[i], 4);
    std::memcpy(&shade[20 * (size_t)i + 19], &s.tri_light[i], 4);
  }
  // lights: power distribution (path.cpp:39-48), areas (triangle.cpp:77-83),
  // the world's bounding sphere and the envmap tables (infinite lights)
  int nl = s.n_lights();
  std::vector<float> area(std::max(nl, 1), 0.f), func(std::max(nl, 1), 0.f), cdf(nl + 1, 0.f);
  std::vector<float> dev_lights(s.lights);
  for (int l = 0; l < nl; ++l) {
    int type = (int)s.lights[32 * (size_t)l];
    if (type < LIGHT_AREA || type > LIGHT_ENVMAP) { delete sc; set_error("unknown light type"); return ORION_ERR_ARG; }
    if (type == LIGHT_AREA) {
      int t = (int)s.lights[32 * (size_t)l + 1];
      if (t < 0 || t >= n) { delete sc; set_error("area light triangle out of range"); return ORION_ERR_ARG; }
    }
  }
  EnvTables env;
  std::string env_err;
  if (!build_env_tables(s, &env, &env_err)) { delete sc; set_error(env_err); return ORION_ERR_ARG; }
  float wc[3], wr;
  world_sphere(s, wc, &wr);
  std::vector<float> power;
  light_powers(s, env, wr, &power);
  for (int l = 0, e = 0; l < nl; ++l) {
    const float *L = &s.lights[32 * (size_t)l];
    if ((int)L[0] == LIGHT_AREA) {
      const float *q = &s.tri_pn[18 * (size_t)(int)L[1]];
      float e1x = q[3] - q[0], e1y = q[4] - q[1], e1z = q[5] - q[2];
      float e2x = q[6] - q[0], e2y = q[7] - q[1], e2z = q[8] - q[2];
      float cx = (e1y * e2z) - (e1z * e2y), cy = (e1z * e2x) - (e1x * e2z), cz = (e1x * e2y) - (e1y * e2x);
      area[l] = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
    }
    if ((int)L[0] == LIGHT_ENVMAP) dev_lights[32 * (size_t)l + 1] = (float)e++;  // device: envmap ordinal
    const float *p = &power[3 * (size_t)l];
    func[l] = 0.212671f * p[0] + 0.715160f * p[1] + 0.072169f * p[2];  // Spectrum::intensity
  }
  float funcInt = 0;
  if (nl > 0) {  // Distribution1D (sampling.h:76-90)
    cdf[0] = 0;
    for (int i = 1; i < nl + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / nl;
    funcInt = cdf[nl];
    if (funcInt == 0) {
      for (int i = 1; i < nl + 1; ++i) cdf[i] = float(i) / float(nl);
    } else {
      for (int i = 1; i < nl + 1; ++i) cdf[i] /= funcInt;
    }
  }
  // upload
  auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
    int rc = b.ensure(bytes);
    if (rc) return rc;
    if (bytes) {
      hipError_t e = hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) { set_error(std::string("upload: ") + hipGetErrorString(e)); return ORION_ERR_HIP; }
    }
    return ORION_OK;
  };
  int rc;
  if ((rc = up(sc->nodes, sc->bvh.nodes.data(), sc->bvh.nodes.size() * sizeof(BvhNode)))) return fail(rc);
  // 16 bytes past the last triangle: the traversal fetches four float4s for a
  // triangle's three (one fetch shape for nodes and triangles)
  if ((rc = sc->tris.ensure(sc->bvh.tris.size() * sizeof(BvhTri) + 16))) return fail(rc);
  if ((rc = up(sc->tris, sc->bvh.tris.data(), sc->bvh.tris.size() * sizeof(BvhTri)))) return fail(rc);
  if ((rc = up(sc->shade, shade.data(), shade.size() * 4))) return fail(rc);
  if ((rc = up(sc->mats, s.mats.data(), s.mats.size() * 4))) return fail(rc);
  if ((rc = up(sc->lights, dev_lights.data(), dev_lights.size() * 4))) return fail(rc);
  if ((rc = up(sc->env_tex, env.tex.data(), env.tex.size() * 4))) return fail(rc);
  if ((rc = up(sc->env_dist, env.dist.data(), env.dist.size() * 4))) return fail(rc);
  if ((rc = up(sc->env_meta, env.meta.data(), env.meta.size() * 4))) return fail(rc);
  if ((rc = up(sc->light_cdf, cdf.data(), cdf.size() * 4))) return fail(rc);
  if ((rc = up(sc->light_func, func.data(), func.size() * 4))) return fail(rc);
  if ((rc = up(sc->light_area, area.data(), area.size() * 4))) return fail(rc);
  const std::vector<uint32_t> &sob = sobol_byte_tables();
  if ((rc = up(sc->sobol, sob.data(), sob.size() * 4))) return fail(rc);
  if ((rc = sc->film.ensure((size_t)W * H * 3 * 4))) return fail(rc);
  if ((rc = sc->count.ensure((size_t)W * H * 4))) return fail(rc);
  if ((rc = sc->ctrl.ensure(64 * 4))) return fail(rc);
  if ((rc = sc->stats.ensure(kStatBytes))) return fail(rc);
  hipError_t e1 = hipMemset(sc->film.ptr, 0, (size_t)W * H * 3 * 4);
  hipError_t e2 = hipMemset(sc->count.ptr, 0, (size_t)W * H * 4);
  hipError_t e3 = hipMemset(sc->ctrl.ptr, 0, 64 * 4);
  hipError_t e4 = hipDeviceSynchronize();
  if (e1 || e2 || e3 || e4) { set_error("scene init memset failed"); return fail(ORION_ERR_HIP); }
  auto t2 = std::chrono::steady_clock::now();
  DScene &d = sc->d;
  d.nodes = (const float4 *)sc->nodes.ptr;
  d.tris = (const float4 *)sc->tris.ptr;
  d.shade = (const float4 *)sc->shade.ptr;
  d.mats = (const float *)sc->mats.ptr;
  d.lights = (const float *)sc->lights.ptr;
  d.light_cdf = (const float *)sc->light_cdf.ptr;
  d.light_func = (const float *)sc->light_func.ptr;
  d.light_area = (const float *)sc->light_area.ptr;
  d.sobol = (const uint32_t *)sc->sobol.ptr;
  d.light_func_int = funcInt;
  d.env_tex = (const float *)sc->env_tex.ptr;
  d.env_dist = (const float *)sc->env_dist.ptr;
  d.env_meta = (const int *)sc->env_meta.ptr;
  d.n_env = (int)(env.meta.size() / 5);
  d.world_c[0] = wc[0]; d.world_c[1] = wc[1]; d.world_c[2] = wc[2];
  d.world_r = wr;
  d.n_lights = nl;
  d.n_prims = n;
  d.stack_bound = sc->bvh.stack_bound;
  std::memcpy(d.cam, s.camera, sizeof(d.cam));
  sc->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  sc->upload_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  c->children.fetch_add(1);
  *out = sc;
  return ORION_OK;
}

int orion_scene_destroy(orion_scene s) {
  if (!s) return ORION_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->ctx->runner) forget_scene(s->ctx->runner, s);
  s->ctx->children.fetch_sub(1);
  delete s;
  return ORION_OK;
}

int orion_scene_get_stats(orion_scene s, orion_scene_stats *st) {
  if (!s || !st) { set_error("null argument"); return ORION_ERR_ARG; }
  st->n_nodes = (int64_t)s->bvh.nodes.size();
  st->n_tris = (int64_t)s->bvh.tris.size();
  st->node_bytes = st->n_nodes * (int64_t)sizeof(BvhNode);
  st->tri_bytes = st->n_tris * (int64_t)sizeof(BvhTri);
  st->shade_bytes = st->n_tris * 80;
  st->bvh_depth = s->bvh.depth;
  st->node_format = 4;
  st->build_ms = s->build_ms;
  st->upload_ms = s->upload_ms;
  st->stack_bound = s->d.stack_bound;
  HIP_TRY(hipSetDevice(s->ctx->device));
  st->trace_waves_per_simd = trace_waves_per_simd(kDefaultRing);
  return ORION_OK;
}

// Persistent intersection launch with the scene's own stack spill area.
static int intersect_launch(orion_scene s, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int *occ, bool any) {
  orion_ctx c = s->ctx;
  int rc = s->spill.ensure(
      trace_spill_ints(s->d.stack_bound, kDefaultRing, trace_grid(1, n, false, kDefaultChunk)) * 4 + 4);
  if (rc) return rc;
  // the scene's statistics area counts dropped stack entries of this call
  HIP_TRY(hipMemsetAsync(s->stats.ptr, 0, kStatBytes, c->stream));
  launch_intersect_batch(s->d, n, rays, t, prim, u, v, tie, occ, any, (int *)s->spill.ptr,
                         std::max(0, s->d.stack_bound - kDefaultRing), (unsigned long long *)s->stats.ptr,
                         c->stream);
  HIP_TRY(hipGetLastError());
  return ORION_OK;
}

// After an intersection launch has completed: fails the call when the
// traversal dropped stack entries.
static int intersect_check(orion_scene s) {
  unsigned long long h[S_COUNT];
  int rc = read_striped_stats(s->stats.ptr, h);
  if (rc) return rc;
  return check_overflow(h[S_OVERFLOW]);
}

int orion_intersect_closest(orion_scene s, int n, const float *rays, float *t, int *prim, float *u, float *v,
                            int *tie, int on_device) {
  if (!s || n < 0 || !rays || !t || !prim) { set_error("bad argument"); return ORION_ERR_ARG; }
  if (n == 0) return ORION_OK;
  orion_ctx c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  if (on_device) {
    int rc = intersect_launch(s, n, rays, t, prim, u, v, tie, nullptr, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return intersect_check(s);
  }
  DevBuf buf;
  int rc = buf.ensure((size_t)n * (8 + 5) * 4);
  if (rc) return rc;
  float *dr = (float *)buf.ptr;
  float *dt = dr + 8 * (size_t)n, *du = dt + n, *dv = du + n;
  int *dp = (int *)(dv + n), *dtie = dp + n;
  HIP_TRY(hipMemcpyAsync(dr, rays, (size_t)n * 8 * 4, hipMemcpyHostToDevice, c->stream));
  if ((rc = intersect_launch(s, n, dr, dt, dp, du, dv, dtie, nullptr, false))) return rc;
  HIP_TRY(hipMemcpyAsync(t, dt, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(prim, dp, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  if (u) HIP_TRY(hipMemcpyAsync(u, du, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  if (v) HIP_TRY(hipMemcpyAsync(v, dv, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  if (tie) HIP_TRY(hipMemcpyAsync(tie, dtie, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return intersect_check(s);
}

int orion_intersect_any(orion_scene s, int n, const float *rays, int *occluded, int on_device) {
  if (!s || n < 0 || !rays || !occluded) { set_error("bad argument"); return ORION_ERR_ARG; }
  if (n == 0) return ORION_OK;
  orion_ctx c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  if (on_device) {
    int rc = intersect_launch(s, n, rays, nullptr, nullptr, nullptr, nullptr, nullptr, occluded, true);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return intersect_check(s);
  }
  DevBuf buf;
  int rc = buf.ensure((size_t)n * 9 * 4);
  if (rc) return rc;
  float *dr = (float *)buf.ptr;
  int *docc = (int *)(dr + 8 * (size_t)n);
  HIP_TRY(hipMemcpyAsync(dr, rays, (size_t)n * 8 * 4, hipMemcpyHostToDevice, c->stream));
  if ((rc = intersect_launch(s, n, dr, nullptr, nullptr, nullptr, nullptr, nullptr, docc, true))) return rc;
  HIP_TRY(hipMemcpyAsync(occluded, docc, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return intersect_check(s);
}

int orion_intersect_closest_noalloc(orion_scene s, int n, const float *rays, float *t, int *prim, float *u,
                                    float *v, int *tie) {
  if (!s || n < 0 || !rays || !t || !prim) { set_error("bad argument"); return ORION_ERR_ARG; }
  if (n == 0) return ORION_OK;
  orion_ctx c = s->ctx;
  const int cap = std::max(0, s->d.stack_bound - kDefaultRing);
  if (cap > kNoAllocSpill) {
    set_error("scene's traversal stack bound " + std::to_string(s->d.stack_bound) +
              " exceeds the allocation-free entry point's maximum " + std::to_string(kNoAllocSpill + kDefaultRing));
    return ORION_ERR_UNSUPPORTED;
  }
  HIP_TRY(hipSetDevice(c->device));
  int *dropped = c->scratch + kNoAllocSpillInts;
  // the counter starts at zero (a kernel, not a memset: callers may capture)
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c->stream, dropped, 64);
  const int blocks = (int)std::min<int64_t>(kNoAllocBlocks, ((int64_t)n + kDefaultChunk - 1) / kDefaultChunk);
  launch_intersect_fixed(s->d, n, rays, t, prim, u, v, tie, c->scratch, cap, blocks, (unsigned int *)dropped,
                         c->stream);
  HIP_TRY(hipGetLastError());
  unsigned int ndrop = 0;
  HIP_TRY(hipMemcpyAsync(&ndrop, dropped, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return check_overflow(ndrop);
}

int orion_film_clear(orion_scene s) {
  if (!s) { set_error("null scene"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(s->ctx->device));
  HIP_TRY(hipMemsetAsync(s->film.ptr, 0, (size_t)s->W * s->H * 3 * 4, s->ctx->stream));
  HIP_TRY(hipMemsetAsync(s->count.ptr, 0, (size_t)s->W * s->H * 4, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return ORION_OK;
}

}  // extern "C"

namespace {
// RenderTarget::getAveraged: Spectrum / int, a true division per channel
__global__ void k_film_average(const float *sum, const int *count, float *avg, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float c = (float)count[i];
    avg[3 * (size_t)i] = sum[3 * (size_t)i] / c;
    avg[3 * (size_t)i + 1] = sum[3 * (size_t)i + 1] / c;
    avg[3 * (size_t)i + 2] = sum[3 * (size_t)i + 2] / c;
  }
}
}  // namespace

extern "C" {

int orion_film_read_averaged(orion_scene s, float *rgb) {
  if (!s || !rgb) { set_error("null argument"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(s->ctx->device));
  const int n = s->W * s->H;
  int rc = s->film_avg.ensure((size_t)n * 3 * 4);
  if (rc) return rc;
  int blocks = std::min((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_film_average, dim3(blocks), dim3(256), 0, s->ctx->stream, (const float *)s->film.ptr,
                     (const int *)s->count.ptr, (float *)s->film_avg.ptr, n);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(rgb, s->film_avg.ptr, (size_t)n * 3 * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return ORION_OK;
}

int orion_film_read(orion_scene s, float *rgb, int *count) {
  if (!s || !rgb || !count) { set_error("null argument"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(s->ctx->device));
  HIP_TRY(hipMemcpyAsync(rgb, s->film.ptr, (size_t)s->W * s->H * 3 * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipMemcpyAsync(count, s->count.ptr, (size_t)s->W * s->H * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return ORION_OK;
}

}  // extern "C"

orion_scene_s::~orion_scene_s() { orion_amd::destroy_sppm(sppm); }
