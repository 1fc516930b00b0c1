// Wave runner shared by the plain renderer, the batched launch and the
// serving loop; the autotuner.
//
// A wave = a device job table + per-wave sizes in the control block +
// raygen, (maxDepth + 1) x [closest trace, any trace, queue reset, shade],
// film.  Every launch argument is fixed for a given (buffers, depth, stack
// variant, grid), so the sequence is captured once as a hipGraph and
// replayed per wave after the job table is copied in.  Queue counters are
// reset by kernels, not hipMemsetAsync, inside the captured region.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/orion_amd.h"
#include "host_scene.h"
#include "kernels.h"
#include "runtime.h"

namespace orion_amd {

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
      return ORION_ERR_HIP;                                                   \
    }                                                                         \
  } while (0)

// Paths per wave: the traversal launches are latency-bound (a launch lasts
// at least as long as its slowest rays), so the wave is made large enough
// that each bounce's launch is dominated by throughput (64 samples per pixel
// of a 1024x1024 film; 208 B of wave state per path = 14 GB of MI355X's 288).
// The server caps its waves at half that: a request that arrives waits for
// the running wave, so a smaller wave bounds its wait (measured: 2^26-path
// served waves raised the median latency from 0.10 to 0.15 s).
constexpr int kDefaultWavePaths = 1 << 26;
constexpr int kServeMaxWavePaths = 1 << 25;

// Per-launch HIP events recorded on the launch stream (kernel timing).
struct LaunchTimer {
  bool on = false;
  hipStream_t st = nullptr;
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;  // 0 closest, 1 shadow, 2 shade, 3 closest + shadow
  int mark(int k) {
    if (!on) return ORION_OK;
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(e, st));
    ev.push_back(e);
    kind.push_back(k);
    return ORION_OK;
  }
  int collect(orion_render_stats *s) {
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      if (!s) continue;
      if (kind[i] == 0) { s->trace_closest_ms += ms; s->trace_closest_launches++; }
      else if (kind[i] == 1) { s->trace_shadow_ms += ms; s->trace_shadow_launches++; }
      else if (kind[i] == 3) { s->trace_pair_ms += ms; s->trace_pair_launches++; }
      else { s->shade_ms += ms; s->shade_launches++; }
    }
    return ORION_OK;
  }
  ~LaunchTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

struct GraphKey {
  int depth, ring, multi, chunk, cap, spill_cap, seg, nseg, split;
  void *mem, *spill;
  bool operator==(const GraphKey &o) const {
    return depth == o.depth && ring == o.ring && multi == o.multi && chunk == o.chunk && cap == o.cap &&
           spill_cap == o.spill_cap && seg == o.seg && nseg == o.nseg && split == o.split && mem == o.mem &&
           spill == o.spill;
  }
};

// Traversal grids of the one-job path, planned from the previous call with
// the same segment size, depth and chunk: the longest segment of each
// bounce's closest-hit and any-hit queues there (the traversal records
// them).  A bounce's launch then covers only the chunk positions such
// segments need, with a margin; its waves take every cps-th chunk of their
// segment, so a longer segment than planned is still traced in full.  Deep
// bounces hold a few thousand paths in thousands of segments: a full grid
// dispatches ~786k one-wave blocks for them.
struct GridPlan {
  const void *scene = nullptr;
  int seg = 0, depth = -1, chunk = 0;
  std::vector<int> longest[2];  // per bounce: closest-hit, any-hit
  // chunk positions per segment for bounce b's queue q (0: every chunk of a
  // full segment)
  int cps(int b, int q, int seg_entries, int chunk) const {
    if (b >= (int)longest[q].size()) return 0;
    const int64_t need = (int64_t)longest[q][b] + longest[q][b] / 4 + 32;
    const int64_t full = (seg_entries + chunk - 1) / chunk;
    const int64_t c = (need + chunk - 1) / chunk;
    return c >= full ? 0 : (int)std::max<int64_t>(1, c);
  }
};

struct WaveRunner {
  WaveBuffers wb;
  DevBuf jobs_dev, ctrl, stats, spill, segmax;
  GridPlan plan;
  bool planned = false;   // the current call's one-job waves use `plan`
  // a plan the caller made for the next wave (the server: from the previous
  // wave when it held the same jobs); used when its segment size, depth and
  // chunk match the wave's
  const GridPlan *wave_plan = nullptr;
  bool one_wave = false;  // ... or one wave per queue segment (orion_tuning.trace_plan 2)
  int spill_cap = 0;
  Job *h_jobs = nullptr;
  int *h_ctrl = nullptr;
  int max_jobs = 0;
  // segment -> job index of multi-job waves (pinned staging + device copy)
  DevBuf segjob_dev;
  int *h_segjob = nullptr;
  DevBuf films;  // batched launches: the jobs' films
  int max_segs = 0;
  hipEvent_t staged = nullptr;
  // the any-hit traversal's stream: a bounce's closest-hit and any-hit
  // launches run side by side (fork / join through two events)
  hipStream_t st_any = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;

  ~WaveRunner() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    if (h_jobs) (void)hipHostFree(h_jobs);
    if (h_ctrl) (void)hipHostFree(h_ctrl);
    if (h_segjob) (void)hipHostFree(h_segjob);
    if (staged) (void)hipEventDestroy(staged);
    if (st_any) (void)hipStreamDestroy(st_any);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
  }
  int ensure_segjob(int nseg) {
    if (nseg <= max_segs) return ORION_OK;
    int cap = std::max(nseg, 2 * max_segs);
    if (h_segjob) {
      HIP_TRY(hipEventSynchronize(staged));
      (void)hipHostFree(h_segjob);
      h_segjob = nullptr;
      max_segs = 0;
    }
    // job of each segment, then the segment order (seg_perm)
    HIP_TRY(hipHostMalloc((void **)&h_segjob, 2 * sizeof(int) * cap, hipHostMallocDefault));
    int rc = segjob_dev.ensure(2 * sizeof(int) * cap);
    if (rc) return rc;
    max_segs = cap;
    drop_graphs();
    return ORION_OK;
  }
  int init(int njobs) {
    if (njobs <= max_jobs && staged) return ORION_OK;
    int cap = std::max(njobs, std::max(64, max_jobs * 2));
    if (h_jobs) {
      HIP_TRY(hipEventSynchronize(staged));
      (void)hipHostFree(h_jobs);
      h_jobs = nullptr;
      max_jobs = 0;
    }
    HIP_TRY(hipHostMalloc((void **)&h_jobs, sizeof(Job) * cap, hipHostMallocDefault));
    if (!h_ctrl) HIP_TRY(hipHostMalloc((void **)&h_ctrl, 64, hipHostMallocDefault));
    int rc = jobs_dev.ensure(sizeof(Job) * cap);
    if (rc) return rc;
    if ((rc = ctrl.ensure(64 * 4))) return rc;
    if ((rc = stats.ensure(kStatBytes))) return rc;
    if ((rc = segmax.ensure(kSegMaxBytes))) return rc;
    HIP_TRY(hipMemset(ctrl.ptr, 0, 64 * 4));
    if (!staged) HIP_TRY(hipEventCreateWithFlags(&staged, hipEventDisableTiming));
    if (!st_any) {
      HIP_TRY(hipStreamCreateWithFlags(&st_any, hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    max_jobs = cap;
    drop_graphs();
    return ORION_OK;
  }
  void drop_graphs() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  int ensure_cap(int paths) {
    if (paths <= wb.cap && wb.mem.ptr) return ORION_OK;
    drop_graphs();
    return wb.ensure(paths);
  }
  // traversal stack spill area for a wave of nseg segments of seg paths
  int ensure_spill(int stack_bound, int ring, int chunk, int seg, int nseg) {
    size_t need = trace_spill_ints(stack_bound, ring, trace_pair_grid(nseg, seg, chunk)) * 4 + 4;
    if (need > spill.bytes) {
      drop_graphs();
      int rc = spill.ensure(need);
      if (rc) return rc;
    }
    spill_cap = std::max(0, stack_bound - ring_variant(ring));
    return ORION_OK;
  }
  // Wave state and traversal spill area for waves of up to `paths` paths in
  // segments of `seg`; ORION_ERR_OOM leaves the runner usable (a caller
  // retries with a smaller wave).
  int reserve(int paths, int seg, int stack_bound, int ring, int chunk) {
    int rc = ensure_cap(paths);
    if (rc) return rc;
    return ensure_spill(stack_bound, ring, chunk, seg, std::max(1, (paths + seg - 1) / seg));
  }
  Wave wave() {
    Wave w = wb.wave;
    w.ctrl = (int *)ctrl.ptr;
    w.stats = (unsigned long long *)stats.ptr;
    w.segmax = (unsigned long long *)segmax.ptr;
    w.spill = (int *)spill.ptr;
    w.spill_cap = spill_cap;
    return w;
  }
  int reset_stats(hipStream_t st) {
    HIP_TRY(hipMemsetAsync(stats.ptr, 0, kStatBytes, st));
    HIP_TRY(hipMemsetAsync(segmax.ptr, 0, kSegMaxBytes, st));
    return ORION_OK;
  }
  // After a one-job call (stream synchronised): the longest segments its
  // waves recorded become the plan for the next call of that shape.
  // The longest queue segments the last waves recorded (stream synchronised)
  // as a plan for waves of this segment size, depth and chunk.
  int read_plan(GridPlan *gp, int seg, int depth, int chunk) {
    std::vector<unsigned long long> h((size_t)kSegMaxBounces * 2 * kStatStripes);
    HIP_TRY(hipMemcpy(h.data(), segmax.ptr, kSegMaxBytes, hipMemcpyDeviceToHost));
    const int nb = std::min(depth + 1, kSegMaxBounces);
    for (int q = 0; q < 2; ++q) {
      gp->longest[q].assign(nb, 0);
      for (int b = 0; b < nb; ++b) {
        unsigned long long m = 0;
        for (int k = 0; k < kStatStripes; ++k) m = std::max(m, h[((size_t)b * 2 + q) * kStatStripes + k]);
        gp->longest[q][b] = (int)std::min<unsigned long long>(m, 1u << 30);
      }
    }
    gp->seg = seg;
    gp->depth = depth;
    gp->chunk = chunk;
    return ORION_OK;
  }
  int record_plan(const void *scene, int seg, int depth, int chunk) {
    int rc = read_plan(&plan, seg, depth, chunk);
    plan.scene = scene;
    return rc;
  }
  int read_stats(orion_render_stats *o) {
    unsigned long long h[S_COUNT];
    int rc = read_striped_stats(stats.ptr, h);
    if (rc) return rc;
    if ((rc = check_overflow(h[S_OVERFLOW]))) return rc;
    if (!o) return ORION_OK;
    o->paths += (int64_t)h[S_PATHS];
    o->closest_rays += (int64_t)h[S_CLOSEST];
    o->shadow_rays += (int64_t)h[S_SHADOW];
    o->ref_equiv_rays += (int64_t)(h[S_CLOSEST] + h[S_SHADOW] + h[S_REFEQ]);
    o->node_visits_closest += (int64_t)h[S_NODES_C];
    o->tri_tests_closest += (int64_t)h[S_TRIS_C];
    o->node_visits_shadow += (int64_t)h[S_NODES_S];
    o->tri_tests_shadow += (int64_t)h[S_TRIS_S];
    o->stack_overflows += (int64_t)h[S_OVERFLOW];
    o->shaded_segments += (int64_t)h[S_SHADED];
    return ORION_OK;
  }

  // A bounce traces its closest-hit queue on the wave's stream and its
  // any-hit queue on st_any at the same time: separate kernels hold 68 and
  // 65 VGPRs (7 waves/SIMD) where one launch of both parts held 74 (6), and
  // the any-hit blocks still fill the GPU while the closest-hit part drains
  // (measured 5% faster traversal than the single paired launch, 8% than two
  // launches in a row).  The any-hit launch takes the spill area's second
  // part.  `split` runs the two launches in a row on one stream.
  void enqueue_wave(const Job *dj, const Wave &w, int depth, int ring, bool multi, int chunk, bool split,
                    hipStream_t st, LaunchTimer *tm, bool use_plan) {
    launch_raygen(dj, w, multi, st);
    Wave wa_base = w;
    wa_base.spill = w.spill + (size_t)w.spill_cap * trace_grid(w.nseg, w.seg, false, chunk) * kTraceBlock;
    // the grid plan: the caller's for this wave, else (plain renders) the
    // previous call's on the same scene
    auto fits = [&](const GridPlan &g) { return g.seg == w.seg && g.depth == depth && g.chunk == chunk; };
    const GridPlan *gp = nullptr;
    if (use_plan && wave_plan && fits(*wave_plan)) gp = wave_plan;
    else if (use_plan && planned && !multi && fits(plan)) gp = &plan;
    for (int it = 0; it <= depth; ++it) {
      Wave wi = w;  // bounce it traces the ray lengths of counter set it % 3
      wi.cnt_ray = w.cnt + (3 * (it % 3) + 1) * w.cnt_stride;
      wi.cnt_shadow = w.cnt + (3 * (it % 3) + 2) * w.cnt_stride;
      // every wave records each bounce's longest segments for later plans
      const int rec_c = it < kSegMaxBounces ? 2 * it : -1;
      const int rec_a = rec_c >= 0 ? rec_c + 1 : -1;
      int cps_c = gp ? gp->cps(it, 0, w.seg, trace_chunk(chunk)) : 0;
      int cps_a = gp ? gp->cps(it, 1, 2 * w.seg, trace_chunk(chunk)) : 0;
      if (one_wave && !multi) cps_c = cps_a = 1;
      if (it == 0) {
        // camera rays only: no any-hit rays exist before the first shading
        // (an empty any-hit launch of the full grid had cost 1 ms beside
        // this one at C2)
        if (tm) tm->mark(split ? 0 : 3);
        launch_trace(dj, wi, false, ring, chunk, multi, st, cps_c, rec_c);
        if (tm) tm->mark(-1);
      } else if (split) {
        if (tm) tm->mark(0);
        launch_trace(dj, wi, false, ring, chunk, multi, st, cps_c, rec_c);
        if (tm) { tm->mark(-1); tm->mark(1); }
        launch_trace(dj, wi, true, ring, chunk, multi, st, cps_a, rec_a);
        if (tm) tm->mark(-1);
      } else {
        if (tm) tm->mark(3);
        Wave wa = wi;
        wa.spill = wa_base.spill;
        (void)hipEventRecord(ev_fork, st);
        (void)hipStreamWaitEvent(st_any, ev_fork, 0);
        launch_trace(dj, wi, false, ring, chunk, multi, st, cps_c, rec_c);
        launch_trace(dj, wa, true, ring, chunk, multi, st_any, cps_a, rec_a);
        (void)hipEventRecord(ev_join, st_any);
        (void)hipStreamWaitEvent(st, ev_join, 0);
        if (tm) tm->mark(-1);
      }
      if (tm) tm->mark(2);
      launch_shade(dj, wi, it, multi, st);
      if (tm) tm->mark(-1);
    }
    launch_film(dj, w, std::max(wb.cap, 1 << 20), multi, st);
  }

  // Shading blocks per segment: enough that a launch has ~kShadeBlocks
  // blocks to balance over the CUs, each part at least one block's worth of
  // paths.
  static int shade_parts(int seg, int nseg) {
    int k = 1;
    while (2 * k * kShadeThreads <= seg && (int64_t)nseg * k < kShadeBlocks) k *= 2;
    return k;
  }
  static constexpr int kShadeBlocks = 8192 * 256 / kShadeThreads;

  // Runs one wave: jobs[0..n) with path0 / pix0 already assigned.
  // seg: queue segment size (0: chosen from npaths); multi-job waves pass
  // the size their job table was aligned to
  int run(hipStream_t st, const Job *jobs, int n, int npaths, int npix, int depth, int ring, int stack_bound,
          bool multi, int chunk, bool split, bool use_graph, LaunchTimer *tm, int seg = 0) {
    int rc = init(n);
    if (rc) return rc;
    if ((rc = ensure_cap(npaths))) return rc;
    int nseg;
    if (seg > 0) nseg = std::max(1, (npaths + seg - 1) / seg);
    else choose_segments(npaths, &seg, &nseg);
    if ((rc = ensure_spill(stack_bound, ring, chunk, seg, nseg))) return rc;
    if ((rc = ensure_segjob(nseg))) return rc;
    // stage the job table and the wave sizes (pinned; the previous copy
    // must have completed before the staging area is rewritten)
    HIP_TRY(hipEventSynchronize(staged));
    std::memcpy(h_jobs, jobs, sizeof(Job) * n);
    // job j owns segments [path0_j / seg, path0_{j+1} / seg)
    for (int j = 0, sg = 0; j < n; ++j) {
      const int end = j + 1 < n ? jobs[j + 1].path0 / seg : nseg;
      for (; sg < end; ++sg) h_segjob[sg] = j;
    }
    // segment order: for each XCD x, the x-th eighth of every job's segments
    int *perm = h_segjob + nseg, np = 0;
    for (int x = 0; x < kXcds; ++x)
      for (int j = 0; j < n; ++j) {
        const int a = jobs[j].path0 / seg, e = j + 1 < n ? jobs[j + 1].path0 / seg : nseg;
        const int len = e - a;
        for (int sg = a + (int)((int64_t)x * len / kXcds); sg < a + (int)((int64_t)(x + 1) * len / kXcds); ++sg)
          perm[np++] = sg;
      }
    if (np != nseg) { set_error("segment order does not cover the wave"); return ORION_ERR_ARG; }
    h_ctrl[0] = npaths;
    h_ctrl[1] = npix;
    h_ctrl[2] = n;
    HIP_TRY(hipMemcpyAsync(jobs_dev.ptr, h_jobs, sizeof(Job) * n, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(segjob_dev.ptr, h_segjob, 2 * sizeof(int) * nseg, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync((int *)ctrl.ptr + C_NPATHS, h_ctrl, 3 * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(staged, st));
    const Job *dj = (const Job *)jobs_dev.ptr;
    Wave w = wave();
    w.seg = seg;
    w.nseg = nseg;
    w.shade_parts = shade_parts(seg, nseg);
    w.seg_job = (const int *)segjob_dev.ptr;
    w.seg_perm = (const int *)segjob_dev.ptr + nseg;
    if (!use_graph || tm) {
      enqueue_wave(dj, w, depth, ring, multi, chunk, split, st, tm, true);
      HIP_TRY(hipGetLastError());
      return ORION_OK;
    }
    GraphKey key{depth, ring, multi ? 1 : 0, chunk, wb.cap, w.spill_cap, seg, nseg, split ? 1 : 0, wb.mem.ptr, spill.ptr};
    hipGraphExec_t exec = nullptr;
    for (auto &g : graphs)
      if (g.first == key) exec = g.second;
    if (!exec) {
      hipGraph_t graph;
      HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue_wave(dj, w, depth, ring, multi, chunk, split, st, nullptr, false);
      HIP_TRY(hipStreamEndCapture(st, &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return ORION_ERR_HIP;
      }
      graphs.push_back({key, exec});
    }
    HIP_TRY(hipGraphLaunch(exec, st));
    return ORION_OK;
  }
};

void destroy_runner(WaveRunner *r) { delete r; }
void forget_scene(WaveRunner *r, const void *scene) {
  if (r->plan.scene == scene) r->plan = GridPlan();
}

// A tuning is checked before anything else: ray slot s of path p is the
// 32-bit index s * N + p (s < 3), so no wave may exceed kMaxWavePathsLimit.
static int check_tuning(const orion_tuning *t) {
  if (!t) return ORION_OK;
  if (t->wave_paths < 0 || t->spp_per_wave < 0 || t->rays_per_wave < 0 || t->lds_ring < 0 || t->queue_segment < 0) {
    set_error("negative tuning value");
    return ORION_ERR_ARG;
  }
  if (t->trace_plan < 0 || t->trace_plan > 2) {
    set_error("trace_plan must be 0, 1 or 2");
    return ORION_ERR_ARG;
  }
  if (t->wave_paths > kMaxWavePathsLimit || t->spp_per_wave > kMaxWavePathsLimit) {
    set_error("wave of more than 2^28 paths: ray-slot indices would overflow 32 bits");
    return ORION_ERR_UNSUPPORTED;
  }
  return ORION_OK;
}

static int ring_of(const orion_tuning *t) { return ring_variant(t ? t->lds_ring : 0); }
// traversal chunk: the tuning's, else by wave (one_job_paths: the paths of a
// one-job wave, 0 for multi-job waves)
static int chunk_of(const orion_tuning *t, int one_job_paths = 0) {
  if (t && t->rays_per_wave > 0) return trace_chunk(t->rays_per_wave);
  return one_job_paths >= kLargeWavePaths ? kLargeWaveChunk : kDefaultChunk;
}
static bool split_of(const orion_tuning *t) { return t && t->split_trace; }
// queue segment size for a wave of `paths` paths
static int seg_of(const orion_tuning *t, int paths) {
  int seg, nseg;
  choose_segments(paths, &seg, &nseg);
  if (t && t->queue_segment > 0) {
    seg = kSegMin;
    while (seg < kSegMax && seg < t->queue_segment) seg *= 2;
  }
  return seg;
}

static Job make_job(orion_scene s, int W, int H, int maxDepth, float *film, int *count) {
  Job J;
  std::memset(&J, 0, sizeof(J));
  J.sc = s->d;
  // the job's camera: the scene's view with this film size (camera.h:31)
  J.sc.cam[13] = (float)W / H;
  J.sc.cam[14] = (float)W;
  J.sc.cam[15] = (float)H;
  J.W = W;
  J.H = H;
  J.npix = W * H;
  J.maxDepth = maxDepth;
  J.film = film;
  J.count = count;
  return J;
}


}  // namespace orion_amd

using namespace orion_amd;

// ======================================================================= plain renderer
extern "C" int orion_render_path(orion_scene s, int k0, int k1, int maxDepth, const orion_tuning *tune,
                                 orion_render_stats *stats) {
  if (int rc = check_tuning(tune)) return rc;
  if (!s) { set_error("null scene"); return ORION_ERR_ARG; }
  if (k0 < 0 || k1 < k0 || maxDepth < 0) { set_error("bad sample range / depth"); return ORION_ERR_ARG; }
  const int npix = s->W * s->H;
  if ((uint64_t)k1 * (uint64_t)npix > 0xffffffffull) {
    set_error("sample index k * W * H must stay below 2^32");
    return ORION_ERR_UNSUPPORTED;
  }
  if (npix > kMaxWavePathsLimit) { set_error("film of more than 2^28 pixels"); return ORION_ERR_UNSUPPORTED; }
  if (maxDepth > 60) { set_error("maxDepth above 60 exceeds the Sobol dimension budget"); return ORION_ERR_UNSUPPORTED; }
  orion_ctx c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  if (!c->runner) c->runner = new WaveRunner();
  WaveRunner &R = *c->runner;
  const int wave_paths = tune && tune->wave_paths > 0 ? tune->wave_paths : kDefaultWavePaths;
  int S = tune && tune->spp_per_wave > 0 ? tune->spp_per_wave : std::max(1, wave_paths / npix);
  S = std::min({S, std::max(1, k1 - k0), kMaxWavePathsLimit / npix});
  int rc;
  if ((rc = R.init(1))) return rc;
  const int ring = ring_of(tune);
  int chunk = chunk_of(tune, S * npix);
  // a wave that does not fit in HBM is halved until it does (down to one
  // sample pass over the film)
  for (;;) {
    chunk = chunk_of(tune, S * npix);
    rc = R.reserve(S * npix, seg_of(tune, S * npix), s->d.stack_bound, ring, chunk);
    if (rc != ORION_ERR_OOM || S == 1) break;
    S = (S + 1) / 2;
  }
  if (rc) return rc;
  if ((rc = R.reset_stats(c->stream))) return rc;
  const bool graph = tune && tune->use_graph;
  const int mode = tune ? tune->trace_plan : 0;
  R.planned = mode == 0 && R.plan.scene == (const void *)s;
  R.one_wave = mode == 2;
  LaunchTimer tm;
  tm.on = tune && tune->profile;
  tm.st = c->stream;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int k = k0; k < k1; k += S) {
    int Sk = std::min(S, k1 - k);
    Job J = make_job(s, s->W, s->H, maxDepth, (float *)s->film.ptr, (int *)s->count.ptr);
    J.S = Sk;
    J.k0 = (uint32_t)k;
    rc = R.run(c->stream, &J, 1, Sk * npix, npix, maxDepth, ring, s->d.stack_bound, false, chunk, split_of(tune),
               graph, tm.on ? &tm : nullptr, seg_of(tune, Sk * npix));
    if (rc) { R.planned = R.one_wave = false; return rc; }
  }
  R.planned = R.one_wave = false;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if ((rc = R.record_plan(s, seg_of(tune, S * npix), maxDepth, chunk))) return rc;
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = ms;
  }
  if ((rc = tm.collect(stats))) return rc;
  return R.read_stats(stats);
}

// ======================================================================= batched launch
namespace orion_amd {

struct JobState {
  orion_job_desc d;
  int next_k = 0;
  float *film = nullptr;  // W * H * 3 sums, then W * H counts
  int *count = nullptr;
  DevBuf own;             // the film's memory when the job owns it
};

constexpr int kMaxWavePaths = 1 << 25;

// Picks each active job's next sample slab for one wave.
//   kPack: jobs in the given order, each taking as many samples as still
//     fit in `target` paths (at least one job's full film per wave), so a
//     wave touches few scenes (scene data stays in the caches) and the
//     tail of one job's samples shares its wave with the next job's start.
//   kFair: the wave holds max(target, one sample of every active job)
//     paths, at most kMaxWavePaths; every job first gets one sample, from a
//     rotating start (*rr) so a job left out of a full wave goes first in the
//     next, then the capacity left is split evenly in whole samples.
// Each job's paths start at a multiple of `seg` (queue segments never mix
// jobs).  taken[i] = samples given to act[i].
// The job table of a wave: jobs act[order[k]] with taken[] samples each,
// paths starting at multiples of seg.
static std::vector<Job> emit_wave(std::vector<JobState *> &act, const std::vector<int> &order,
                                  const std::vector<int> &taken, int seg, int *npaths, int *npix, int *depth,
                                  int *stack) {
  std::vector<Job> jobs;
  int64_t p0 = 0;
  int pix = 0;
  *depth = 0;
  *stack = 1;
  for (int i : order) {
    JobState *js = act[i];
    const int np = js->d.width * js->d.height;
    Job J = make_job(js->d.scene, js->d.width, js->d.height, js->d.max_depth, js->film, js->count);
    J.S = taken[i];
    J.k0 = (uint32_t)js->next_k;
    J.path0 = (int)p0;
    J.pix0 = pix;
    p0 = (p0 + (int64_t)J.S * np + seg - 1) / seg * seg;
    pix += np;
    *depth = std::max(*depth, js->d.max_depth);
    *stack = std::max(*stack, js->d.scene->d.stack_bound);
    jobs.push_back(J);
  }
  *npaths = (int)p0;
  *npix = pix;
  return jobs;
}

enum WavePolicy { kPack, kFair };
static std::vector<Job> compose_wave(std::vector<JobState *> &act, int target, int seg, WavePolicy policy, int *rr,
                                     int *npaths, int *npix, int *depth, int *stack, std::vector<int> *taken) {
  const int n_act = (int)act.size();
  taken->assign(n_act, 0);
  auto padded = [&](int64_t paths) { return (paths + seg - 1) / seg * seg; };
  if (policy == kPack) {
    int max_np = 1;
    for (JobState *js : act) max_np = std::max(max_np, js->d.width * js->d.height);
    const int64_t cap = std::max<int64_t>(padded(max_np), target);
    int64_t paths = 0;
    std::vector<int> order;
    for (int i = 0; i < n_act && paths < cap; ++i) {
      JobState *js = act[i];
      const int64_t np = (int64_t)js->d.width * js->d.height;
      int c = (int)std::min<int64_t>(js->d.spp - js->next_k, (cap - paths) / np);
      while (c > 0 && paths + padded(c * np) > cap) --c;
      if (c <= 0) continue;
      (*taken)[i] = c;
      paths += padded(c * np);
      order.push_back(i);
    }
    return emit_wave(act, order, *taken, seg, npaths, npix, depth, stack);
  }
  int64_t sum_np = 0;
  int max_np = 1;
  for (JobState *js : act)
    if (js->d.spp > js->next_k) {
      int np = js->d.width * js->d.height;
      sum_np += padded(np);
      max_np = std::max(max_np, np);
    }
  const int64_t cap = std::max<int64_t>(padded(max_np), std::min<int64_t>(std::max<int64_t>(target, sum_np),
                                                                           std::max(target, kMaxWavePaths)));
  // one sample each, rotating start
  std::vector<int> order;
  int64_t paths = 0;
  const int start = n_act ? *rr % n_act : 0;
  for (int t = 0; t < n_act; ++t) {
    const int i = (start + t) % n_act;
    JobState *js = act[i];
    const int64_t np = (int64_t)js->d.width * js->d.height;
    if (js->d.spp <= js->next_k || paths + padded(np) > cap) continue;
    (*taken)[i] = 1;
    paths += padded(np);
    order.push_back(i);
  }
  if (!order.empty()) *rr = order.back() + 1;
  // the rest of the capacity, evenly
  if (!order.empty()) {
    int64_t left = cap - paths;
    const int64_t share = left / (int64_t)order.size();
    for (int i : order) {
      JobState *js = act[i];
      const int64_t np = (int64_t)js->d.width * js->d.height;
      int extra = (int)std::min<int64_t>(js->d.spp - js->next_k - 1, share / np);
      // the padded slab may not outgrow the share
      while (extra > 0 && padded((1 + extra) * np) - padded(np) > share) --extra;
      (*taken)[i] += extra;
      paths += padded((1 + extra) * np) - padded(np);
    }
  }
  return emit_wave(act, order, *taken, seg, npaths, npix, depth, stack);
}

// compose_wave with the queue segment the wave's size calls for
// (choose_segments): composed at kSegMax first, and again at kSegMin when
// the wave came out under 2^21 paths.  A tuning that fixes the segment
// (fixed = true) keeps `seg`.
static std::vector<Job> compose_sized(std::vector<JobState *> &act, int target, int seg, bool fixed,
                                      WavePolicy policy, int *rr, int *npaths, int *npix, int *depth, int *stack,
                                      std::vector<int> *taken, int *seg_out) {
  const int rr0 = *rr;
  *seg_out = fixed ? seg : kSegMax;
  std::vector<Job> jobs = compose_wave(act, target, *seg_out, policy, rr, npaths, npix, depth, stack, taken);
  int s2, n2;
  choose_segments(*npaths, &s2, &n2);
  if (!fixed && !jobs.empty() && s2 != *seg_out) {
    *rr = rr0;
    *seg_out = s2;
    jobs = compose_wave(act, target, s2, policy, rr, npaths, npix, depth, stack, taken);
  }
  return jobs;
}

static int check_desc(const orion_job_desc &d) {
  if (!d.scene || d.width <= 0 || d.height <= 0 || d.spp < 0 || d.max_depth < 0 || d.max_depth > 60) {
    set_error("bad job description");
    return ORION_ERR_ARG;
  }
  if ((uint64_t)d.spp * (uint64_t)d.width * (uint64_t)d.height > 0xffffffffull) {
    set_error("sample index k * W * H must stay below 2^32");
    return ORION_ERR_UNSUPPORTED;
  }
  if ((int64_t)d.width * d.height > kMaxWavePathsLimit / 2) {
    // one sample pass of the film, padded to a queue segment, must fit a wave
    set_error("film of more than 2^27 pixels");
    return ORION_ERR_UNSUPPORTED;
  }
  return ORION_OK;
}

}  // namespace orion_amd

extern "C" int orion_render_batch(orion_ctx c, int n, const orion_job_desc *descs, float *const *rgb_out,
                                  int *const *count_out, const orion_tuning *tune, orion_render_stats *stats) {
  if (int rc = check_tuning(tune)) return rc;
  if (!c || n < 0 || (n && (!descs || !rgb_out || !count_out))) { set_error("bad argument"); return ORION_ERR_ARG; }
  for (int j = 0; j < n; ++j) {
    int rc = check_desc(descs[j]);
    if (rc) return rc;
    if (descs[j].scene->ctx != c) { set_error("job scene belongs to another context"); return ORION_ERR_ARG; }
  }
  HIP_TRY(hipSetDevice(c->device));
  if (!c->runner) c->runner = new WaveRunner();
  WaveRunner &R = *c->runner;
  // every job's film (sums then counts) in one region of the runner's film
  // buffer, kept between calls
  std::vector<JobState> js(n);
  size_t film_bytes = 0;
  for (int j = 0; j < n; ++j) film_bytes += (size_t)descs[j].width * descs[j].height * 16;
  {
    int rc = R.films.ensure(std::max<size_t>(film_bytes, 16));
    if (rc) return rc;
  }
  HIP_TRY(hipMemsetAsync(R.films.ptr, 0, film_bytes, c->stream));
  size_t off = 0;
  for (int j = 0; j < n; ++j) {
    js[j].d = descs[j];
    size_t np = (size_t)descs[j].width * descs[j].height;
    js[j].film = (float *)((char *)R.films.ptr + off);
    js[j].count = (int *)((char *)R.films.ptr + off + np * 12);
    off += np * 16;
  }
  int target = tune && tune->wave_paths > 0 ? tune->wave_paths : kDefaultWavePaths;
  int rc;
  if ((rc = R.init(std::max(n, 1)))) return rc;
  {
    // the largest wave: the target, or one sample pass of the largest film
    // padded to a segment; halved while it does not fit in HBM
    int max_np = 1, max_stack = 1;
    for (int j = 0; j < n; ++j) {
      max_np = std::max(max_np, descs[j].width * descs[j].height);
      max_stack = std::max(max_stack, descs[j].scene->d.stack_bound);
    }
    const int floor_paths = (max_np + kSegMax - 1) / kSegMax * kSegMax;
    for (;;) {
      const int paths = std::max(target, floor_paths);
      rc = R.reserve(paths, seg_of(tune, paths), max_stack, ring_of(tune), chunk_of(tune));
      if (rc != ORION_ERR_OOM || target <= floor_paths) break;
      target /= 2;
    }
    if (rc) return rc;
  }
  if ((rc = R.reset_stats(c->stream))) return rc;
  const bool graph = tune && tune->use_graph;
  LaunchTimer tm;
  tm.on = tune && tune->profile;
  tm.st = c->stream;
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  std::vector<JobState *> act;
  for (auto &j : js) act.push_back(&j);
  int rr = 0;
  const int seg = seg_of(tune, target);
  for (;;) {
    int npaths, npix, depth, stack;
    std::vector<int> taken;
    int wseg;
    std::vector<Job> jobs = compose_sized(act, target, seg, tune && tune->queue_segment > 0, kPack, &rr, &npaths,
                                          &npix, &depth, &stack, &taken, &wseg);
    if (jobs.empty()) break;
    if ((rc = R.run(c->stream, jobs.data(), (int)jobs.size(), npaths, npix, depth, ring_of(tune), stack,
                    jobs.size() > 1, chunk_of(tune), split_of(tune), graph, tm.on ? &tm : nullptr, wseg)))
      return rc;
    for (size_t i = 0; i < act.size(); ++i) act[i]->next_k += taken[i];
    act.erase(std::remove_if(act.begin(), act.end(), [](JobState *j) { return j->next_k >= j->d.spp; }), act.end());
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  for (int j = 0; j < n; ++j) {
    size_t np = (size_t)descs[j].width * descs[j].height;
    HIP_TRY(hipMemcpyAsync(rgb_out[j], js[j].film, np * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(count_out[j], js[j].count, np * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->ms = ms;
  }
  if ((rc = tm.collect(stats))) return rc;
  return R.read_stats(stats);
}

// ======================================================================= serving
namespace orion_amd {
struct ServerJob {
  uint64_t id;
  JobState st;
  int state = 0;  // 0 queued, 1 active, 2 done, 3 failed
  double t_submit = 0, t_admit = 0, t_done = 0;
  int64_t bytes = 0;
  std::vector<float> rgb;
  std::vector<int> cnt;
  // progressive preview: a peek asks for a copy of the film between waves
  bool peek_req = false;
  uint64_t peek_seq = 0;
  int peek_samples = 0;
  std::vector<float> prgb;
  std::vector<int> pcnt;
};
}  // namespace orion_amd

struct orion_server_s {
  orion_ctx ctx = nullptr;
  hipStream_t stream = nullptr;  // the serving loop's own stream
  int64_t budget = 0, in_use = 0;
  orion_tuning tune{};
  bool has_tune = false;
  orion_amd::WavePolicy policy = orion_amd::kPack;
  bool shortest_first = true;  // kPack order: least remaining work first, else submission order
  bool aged = false;           // ... with remaining work credited for the time waited
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<orion_amd::ServerJob *> queue;
  std::vector<orion_amd::ServerJob *> active;
  std::map<uint64_t, orion_amd::ServerJob *> all;
  uint64_t next_id = 1;
  bool stop = false;
  int error = 0;
  std::string error_msg;
  std::thread th;
  orion_amd::WaveRunner runner;
  std::chrono::steady_clock::time_point t0;
  int64_t jobs_done = 0, waves = 0, paths = 0, closest = 0, shadow = 0, mixed_waves = 0;
  int retiring = 0;  // finished jobs whose films wait in pinned staging for the host copy
  int max_jobs_in_wave = 0;
  int64_t hbm_peak = 0;
  double busy_ms = 0;
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  void loop();
};

void orion_server_s::loop() {
  using namespace orion_amd;
  (void)hipSetDevice(ctx->device);
  const orion_tuning *tn = has_tune ? &tune : nullptr;
  int target = std::min(tn && tn->wave_paths > 0 ? tn->wave_paths : kDefaultWavePaths, kServeMaxWavePaths);
  int rr = 0;
  const int seg = seg_of(tn, target);
  GridPlan wplan;                // per-bounce longest queue segments of the last wave
  std::vector<int64_t> plan_sig;  // that wave's composition (empty: no plan)
  // Finished films go to pinned staging on the stream right after their
  // wave's kernels; the host copies them out (and reports the jobs done)
  // once the next wave is enqueued, so that copy overlaps GPU work.  Two
  // staging buffers alternate by wave: wave k's copies land in one while the
  // host reads wave k - 1's from the other.
  char *stage[2] = {nullptr, nullptr};
  size_t stage_cap[2] = {0, 0};
  int wave_no = 0;
  struct Retire {
    ServerJob *j;
    const char *src;
  };
  std::vector<Retire> pending;
  auto finish_pending = [&] {
    if (pending.empty()) return;
    for (const Retire &r : pending) {
      const size_t np = (size_t)r.j->st.d.width * r.j->st.d.height;
      r.j->rgb.resize(np * 3);
      r.j->cnt.resize(np);
      std::memcpy(r.j->rgb.data(), r.src, np * 12);
      std::memcpy(r.j->cnt.data(), r.src + np * 12, np * 4);
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      for (const Retire &r : pending) {
        r.j->state = 2;
        r.j->t_done = now();
        jobs_done++;
      }
      retiring -= (int)pending.size();
    }
    pending.clear();
    cv_done.notify_all();
  };
  for (;;) {
    std::vector<ServerJob *> admitted;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto ready = [&] { return stop || !queue.empty() || !active.empty(); };
      if (!ready() && !pending.empty()) {  // idle next: report the last wave's jobs first
        lk.unlock();
        finish_pending();
        lk.lock();
      }
      cv.wait(lk, ready);
      if (stop && active.empty() && queue.empty()) {
        lk.unlock();
        finish_pending();
        for (char *b : stage)
          if (b) (void)hipHostFree(b);
        return;
      }
      // admission against the HBM budget (FIFO; one job is always admitted
      // when nothing runs, so an oversized job cannot stall the queue)
      while (!queue.empty()) {
        ServerJob *j = queue.front();
        if (!active.empty() && in_use + j->bytes > budget) break;
        queue.pop_front();
        in_use += j->bytes;
        hbm_peak = std::max(hbm_peak, in_use);
        admitted.push_back(j);
        active.push_back(j);
      }
    }
    for (ServerJob *j : admitted) {
      size_t np = (size_t)j->st.d.width * j->st.d.height;
      int rc = j->st.own.ensure(np * 16);
      if (!rc) {
        j->st.film = (float *)j->st.own.ptr;
        j->st.count = (int *)((char *)j->st.own.ptr + np * 12);
      }
      if (!rc && hipMemsetAsync(j->st.own.ptr, 0, np * 16, stream) != hipSuccess)
        rc = ORION_ERR_HIP;
      std::lock_guard<std::mutex> lk(mu);
      j->t_admit = now();
      if (rc) {
        j->state = 3;
        error = rc;
        error_msg = orion_last_error();
        in_use -= j->bytes;
        active.erase(std::find(active.begin(), active.end(), j));
      } else {
        j->state = 1;
      }
    }
    if (!admitted.empty()) cv_done.notify_all();
    std::vector<JobState *> act;
    std::vector<ServerJob *> owners;
    WavePolicy policy;
    bool srw, age;
    double rate;  // paths/s over the busy time so far
    {
      std::lock_guard<std::mutex> lk(mu);
      policy = this->policy;
      srw = shortest_first;
      age = aged;
      rate = busy_ms > 0 ? (double)paths / (busy_ms * 1e-3) : 0.0;
      for (ServerJob *j : active)
        if (j->state == 1) { act.push_back(&j->st); owners.push_back(j); }
    }
    if (act.empty()) continue;
    int npaths, npix, depth, stack;
    std::vector<int> taken;
    // packed: smallest remaining work first (shortest-remaining-time order)
    std::vector<size_t> idx(act.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    const double tnow = now();
    auto remaining = [&](size_t i) {
      double r = (double)(act[i]->d.spp - act[i]->next_k) * act[i]->d.width * act[i]->d.height;
      if (age) r -= rate * (tnow - owners[i]->t_submit);
      return r;
    };
    if (policy == kPack && srw)
      std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return remaining(a) < remaining(b); });
    std::vector<JobState *> sorted_act;
    std::vector<ServerJob *> sorted_owners;
    for (size_t i : idx) {
      sorted_act.push_back(act[i]);
      sorted_owners.push_back(owners[i]);
    }
    act.swap(sorted_act);
    owners.swap(sorted_owners);
    int wseg;
    std::vector<Job> jobs = compose_sized(act, target, seg, tn && tn->queue_segment > 0, policy, &rr, &npaths, &npix,
                                          &depth, &stack, &taken, &wseg);
    // a wave with the same jobs, sample counts and layout as the previous
    // one traces each bounce on grids sized from that wave's queues (a
    // request's waves repeat until its last one)
    std::vector<int64_t> sig;
    for (const Job &J : jobs) {
      sig.push_back((int64_t)(intptr_t)J.sc.nodes);
      sig.push_back(((int64_t)J.W << 32) | (uint32_t)J.H);
      sig.push_back(((int64_t)J.S << 32) | (uint32_t)J.maxDepth);
      sig.push_back(J.path0);
    }
    sig.push_back(npaths);
    sig.push_back(wseg);
    const bool graph = tn && tn->use_graph;
    runner.wave_plan = !graph && sig == plan_sig ? &wplan : nullptr;
    // the jobs this wave finishes, and their films' place in this wave's staging
    const int sb = wave_no++ & 1;
    std::vector<size_t> stage_off(act.size(), SIZE_MAX);
    size_t stage_need = 0;
    for (size_t i = 0; i < act.size(); ++i)
      if (act[i]->next_k + taken[i] >= act[i]->d.spp) {
        stage_off[i] = stage_need;
        stage_need += (size_t)act[i]->d.width * act[i]->d.height * 16;
      }
    int rc = ORION_OK;
    if (stage_need > stage_cap[sb]) {  // the host finished reading this buffer a wave ago
      if (stage[sb]) (void)hipHostFree(stage[sb]);
      stage[sb] = nullptr;
      stage_cap[sb] = 0;
      if (hipHostMalloc((void **)&stage[sb], stage_need, hipHostMallocDefault) != hipSuccess) {
        stage[sb] = nullptr;
        set_error("pinned staging allocation failed");
        rc = ORION_ERR_OOM;
      } else {
        stage_cap[sb] = stage_need;
      }
    }
    auto w0 = std::chrono::steady_clock::now();
    if (!rc) rc = runner.init(std::max(1, (int)jobs.size()));
    if (!rc) rc = runner.reset_stats(stream);
    if (!rc && !jobs.empty())
      rc = runner.run(stream, jobs.data(), (int)jobs.size(), npaths, npix, depth, ring_of(tn), stack,
                      jobs.size() > 1, chunk_of(tn), split_of(tn), graph, nullptr, wseg);
    runner.wave_plan = nullptr;
    for (size_t i = 0; i < act.size() && !rc; ++i)  // film sums and counts are contiguous
      if (stage_off[i] != SIZE_MAX &&
          hipMemcpyAsync(stage[sb] + stage_off[i], act[i]->film, (size_t)act[i]->d.width * act[i]->d.height * 16,
                         hipMemcpyDeviceToHost, stream) != hipSuccess)
        rc = ORION_ERR_HIP;
    finish_pending();  // the previous wave's films, while this wave runs
    if (!rc && hipStreamSynchronize(stream) != hipSuccess) rc = ORION_ERR_HIP;
    orion_render_stats ws{};
    if (!rc) rc = runner.read_stats(&ws);
    plan_sig.clear();
    if (!rc && !jobs.empty() && !graph && runner.read_plan(&wplan, wseg, depth, chunk_of(tn)) == ORION_OK)
      plan_sig = sig;
    double wms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    std::vector<ServerJob *> finished;
    std::vector<const char *> finished_src;
    for (size_t i = 0; i < act.size(); ++i) {
      act[i]->next_k += taken[i];
      if (act[i]->next_k >= act[i]->d.spp) {
        finished.push_back(owners[i]);
        finished_src.push_back(stage_off[i] != SIZE_MAX ? stage[sb] + stage_off[i] : nullptr);
      }
    }
    // previews of running jobs asked for since the last wave (the stream is
    // idle here, so the films hold whole sample passes)
    std::vector<ServerJob *> peeks;
    {
      std::lock_guard<std::mutex> lk(mu);
      for (ServerJob *j : active)
        if (j->peek_req && j->state == 1 && std::find(finished.begin(), finished.end(), j) == finished.end())
          peeks.push_back(j);
    }
    for (ServerJob *j : peeks) {
      size_t np = (size_t)j->st.d.width * j->st.d.height;
      j->prgb.resize(np * 3);
      j->pcnt.resize(np);
      if (!rc && (hipMemcpy(j->prgb.data(), j->st.film, np * 12, hipMemcpyDeviceToHost) != hipSuccess ||
                  hipMemcpy(j->pcnt.data(), j->st.count, np * 4, hipMemcpyDeviceToHost) != hipSuccess))
        rc = ORION_ERR_HIP;
    }
    for (ServerJob *j : finished) {  // their films are in staging (the wave has completed)
      j->st.own.release();
      j->st.film = nullptr;
      j->st.count = nullptr;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      waves++;
      if (jobs.size() > 1) mixed_waves++;
      max_jobs_in_wave = std::max(max_jobs_in_wave, (int)jobs.size());
      busy_ms += wms;
      paths += ws.paths;
      closest += ws.closest_rays;
      shadow += ws.shadow_rays;
      if (rc) { error = rc; error_msg = orion_last_error(); }
      for (ServerJob *j : peeks) {
        j->peek_samples = j->st.next_k;
        j->peek_req = false;
        j->peek_seq++;
      }
      for (size_t k = 0; k < finished.size(); ++k) {
        ServerJob *j = finished[k];
        in_use -= j->bytes;
        active.erase(std::find(active.begin(), active.end(), j));
        if (rc) {
          j->state = 3;
          j->t_done = now();
          jobs_done++;
        } else {  // reported once the host has copied the film out
          pending.push_back(Retire{j, finished_src[k]});
          retiring++;
        }
      }
    }
    cv_done.notify_all();
  }
}

extern "C" {

int orion_server_create(orion_ctx c, int64_t budget, const orion_tuning *tune, orion_server *out) {
  if (int rc = orion_amd::check_tuning(tune)) return rc;
  if (!c || !out || budget <= 0) { set_error("bad argument"); return ORION_ERR_ARG; }
  HIP_TRY(hipSetDevice(c->device));
  auto *s = new orion_server_s();
  s->ctx = c;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    set_error(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
    return ORION_ERR_HIP;
  }
  s->budget = budget;
  if (tune) { s->tune = *tune; s->has_tune = true; }
  s->t0 = std::chrono::steady_clock::now();
  s->th = std::thread([s] { s->loop(); });
  c->children.fetch_add(1);
  *out = s;
  return ORION_OK;
}

int orion_server_set_policy(orion_server s, int policy) {
  if (!s || policy < ORION_SERVE_PACK || policy > ORION_SERVE_AGED) {
    set_error("bad argument");
    return ORION_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->policy = policy == ORION_SERVE_FAIR ? orion_amd::kFair : orion_amd::kPack;
  s->shortest_first = policy == ORION_SERVE_PACK || policy == ORION_SERVE_AGED;
  s->aged = policy == ORION_SERVE_AGED;
  return ORION_OK;
}

int orion_server_submit(orion_server s, const orion_job_desc *d, uint64_t *id) {
  if (!s || !d || !id) { set_error("bad argument"); return ORION_ERR_ARG; }
  int rc = orion_amd::check_desc(*d);
  if (rc) return rc;
  auto *j = new orion_amd::ServerJob();
  j->st.d = *d;
  j->bytes = (int64_t)d->width * d->height * 16;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) { delete j; set_error("server stopped"); return ORION_ERR_BUSY; }
    j->id = s->next_id++;
    j->t_submit = s->now();
    s->queue.push_back(j);
    s->all[j->id] = j;
    *id = j->id;
  }
  s->cv.notify_one();
  return ORION_OK;
}

int orion_server_poll(orion_server s, uint64_t id, int *done, float *rgb, int *count, double times[3]) {
  if (!s || !done) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) { set_error("unknown job id"); return ORION_ERR_ARG; }
  orion_amd::ServerJob *j = it->second;
  if (j->state == 3) { set_error("job failed: " + s->error_msg); return s->error ? s->error : ORION_ERR_HIP; }
  *done = j->state == 2;
  if (times) { times[0] = j->t_submit; times[1] = j->t_admit; times[2] = j->t_done; }
  if (*done) {
    if (rgb) std::memcpy(rgb, j->rgb.data(), j->rgb.size() * 4);
    if (count) std::memcpy(count, j->cnt.data(), j->cnt.size() * 4);
  }
  return ORION_OK;
}

int orion_server_peek(orion_server s, uint64_t id, float *rgb, int *count, int *samples) {
  if (!s || !samples) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) { set_error("unknown job id"); return ORION_ERR_ARG; }
  orion_amd::ServerJob *j = it->second;
  const size_t np = (size_t)j->st.d.width * j->st.d.height;
  if (j->state == 1) {  // running: wait for the next wave boundary's copy
    uint64_t seq = j->peek_seq;
    j->peek_req = true;
    s->cv_done.wait(lk, [&] { return j->peek_seq != seq || j->state != 1; });
  }
  if (j->state == 3) { set_error("job failed: " + s->error_msg); return s->error ? s->error : ORION_ERR_HIP; }
  if (j->state == 0) {  // queued: nothing rendered yet
    if (rgb) std::memset(rgb, 0, np * 12);
    if (count) std::memset(count, 0, np * 4);
    *samples = 0;
  } else if (j->state == 2) {
    if (rgb) std::memcpy(rgb, j->rgb.data(), np * 12);
    if (count) std::memcpy(count, j->cnt.data(), np * 4);
    *samples = j->st.d.spp;
  } else {
    if (rgb) std::memcpy(rgb, j->prgb.data(), np * 12);
    if (count) std::memcpy(count, j->pcnt.data(), np * 4);
    *samples = j->peek_samples;
  }
  return ORION_OK;
}

int orion_server_drain(orion_server s) {
  if (!s) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queue.empty() && s->active.empty() && s->retiring == 0; });
  if (s->error) { set_error(s->error_msg); return s->error; }
  return ORION_OK;
}

int orion_server_get_stats(orion_server s, orion_server_stats *st) {
  if (!s || !st) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->mu);
  st->jobs_done = s->jobs_done;
  st->waves = s->waves;
  st->paths = s->paths;
  st->closest_rays = s->closest;
  st->shadow_rays = s->shadow;
  st->active_jobs = (int)s->active.size();
  st->queued_jobs = (int)s->queue.size();
  st->hbm_in_use = s->in_use;
  st->busy_ms = s->busy_ms;
  st->mixed_waves = s->mixed_waves;
  st->max_jobs_in_wave = s->max_jobs_in_wave;
  st->hbm_peak = s->hbm_peak;
  return ORION_OK;
}

int orion_server_destroy(orion_server s) {
  if (!s) return ORION_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->stream);
  for (auto &kv : s->all) delete kv.second;
  s->runner.drop_graphs();
  (void)hipStreamDestroy(s->stream);
  s->ctx->children.fetch_sub(1);
  delete s;
  return ORION_OK;
}

// ======================================================================= autotuner
int orion_autotune(orion_ctx c, int n, const orion_job_desc *jobs, orion_tuning *best) {
  if (!c || n <= 0 || !jobs || !best) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::vector<orion_job_desc> trial_jobs(jobs, jobs + n);
  std::vector<std::vector<float>> rgb(n);
  std::vector<std::vector<int>> cnt(n);
  std::vector<float *> rp(n);
  std::vector<int *> cp(n);
  for (int j = 0; j < n; ++j) {
    trial_jobs[j].spp = std::min(trial_jobs[j].spp, 2);
    size_t np = (size_t)trial_jobs[j].width * trial_jobs[j].height;
    rgb[j].resize(np * 3);
    cnt[j].resize(np);
    rp[j] = rgb[j].data();
    cp[j] = cnt[j].data();
  }
  // coordinate descent from the defaults: wave size, then traversal chunk,
  // LDS ring, queue segment, graph, split launches.  A candidate replaces the
  // current choice only when it is at least 2% faster: a cut-down mix often
  // cannot tell two settings apart (every wave size above its path count
  // gives the same single wave), and noise must not pick for the full run.
  double best_ms = 1e30;
  orion_tuning b{};
  b.wave_paths = kDefaultWavePaths;
  auto trial = [&](const orion_tuning &t) -> int {
    orion_render_stats st{};
    // the first call captures graphs and sizes buffers; time the second
    int rc = orion_render_batch(c, n, trial_jobs.data(), rp.data(), cp.data(), &t, &st);
    if (!rc) rc = orion_render_batch(c, n, trial_jobs.data(), rp.data(), cp.data(), &t, &st);
    if (rc) return rc;
    if (st.ms < 0.98 * best_ms) { best_ms = st.ms; b = t; }
    return ORION_OK;
  };
  int rc;
  if ((rc = trial(b))) return rc;
  for (int wp : {1 << 25, 1 << 24, 1 << 23}) {
    orion_tuning t{};
    t.wave_paths = wp;
    if ((rc = trial(t))) return rc;
  }
  // chunks below the default are not tried: the cut-down trial mix (small
  // waves) favours them, full-size waves lose with them (128-entry chunks: a
  // served stream 6% slower)
  for (int chunk : {512}) {
    orion_tuning t = b;
    t.rays_per_wave = chunk;
    if ((rc = trial(t))) return rc;
  }
  for (int ring : {8, 32}) {
    orion_tuning t = b;
    t.lds_ring = ring;
    if ((rc = trial(t))) return rc;
  }
  for (int seg : {1024, 2048, 4096, 8192, 16384}) {
    orion_tuning t = b;
    t.queue_segment = seg;
    if ((rc = trial(t))) return rc;
  }
  {
    orion_tuning t = b;
    t.use_graph = 1;
    if ((rc = trial(t))) return rc;
  }
  {
    orion_tuning t = b;
    t.split_trace = 1;
    if ((rc = trial(t))) return rc;
  }
  if (b.rays_per_wave == 0) b.rays_per_wave = kDefaultChunk;
  if (b.lds_ring == 0) b.lds_ring = kDefaultRing;
  *best = b;  // queue_segment 0: chosen per wave from its size
  return ORION_OK;
}

}  // extern "C"
