/* orion_amd: MI355X-native (gfx950) path-tracing + SPPM hot path for orion.
 *
 * Plain C ABI: opaque handles, plain pointers and sizes, int status
 * (0 = ok, < 0 = error; orion_last_error() returns the message of the last
 * failure on the calling thread).  No exceptions cross this boundary.
 *
 * Where each entry point sits in orion (paths relative to the reference src/):
 *   orion_hscene_load_ori     replaces Parser::Parser + makeRenderOption
 *                             (util/parser.cpp:45-56, 463-471)
 *   orion_scene_upload        replaces Scene::Scene -> EmbreeKernel::build
 *                             (core/scene.cpp:5-15, core/kernel.cpp:63-110)
 *   orion_intersect_closest   replaces EmbreeKernel::intersect -> rtcIntersect
 *                             (core/kernel.cpp:111-174), batched
 *   orion_intersect_any       replaces ShadowTester::unoccluded
 *                             (core/light.h:26-29), batched
 *   orion_render_path         replaces PathTracing::render (integrator/path.cpp:15-37)
 *   orion_render_sppm         replaces SPPM::render (integrator/sppm.cpp:58-322)
 *   orion_film_read           replaces RenderTarget::getAveraged's inputs
 *                             (texture/rendertarget.cpp:22-37): sum + count
 *   orion_film_read_averaged  replaces RenderTarget::getAveraged
 *   orion_film_write          replaces saveAsBmp (texture/texture.cpp:77-137),
 *                             plus FP32 PFM / EXR
 *   orion_batch_* / orion_server_*  no counterpart in orion (batched launch and
 *                             continuous-batching serving of many render jobs)
 *
 * Resolved scene records (float32), shared by every consumer of a scene:
 *   triangle  18 floats  p0 p1 p2 n0 n1 n2, world space (Mesh de-indexed,
 *                        core/triangle.h:16-25, after createTriangleMesh)
 *   tri_mat   material index per triangle; tri_light area-light index or -1
 *   material  ORION_MAT_FLOATS (24) floats  [0] type: 0 matte, 1 metal,
 *             2 glass, 3 mirror, 4 plastic, 5 uber
 *             matte:   [1..3] Kd  [4] sigma (0: Lambert, else Oren-Nayar)
 *             metal:   [1..3] eta  [4..6] k  [7] roughness  [8] alpha
 *             glass:   [1..3] Kr  [4..6] Kt  [7] roughness  [8] alpha  [9] eta
 *             mirror:  [1..3] Kr
 *             plastic: [1..3] Kd  [4..6] Ks  [7] roughness  [8] alpha
 *             uber:    [1..3] Kd  [4..6] Ks  [7..9] Kr  [10..12] Kt
 *                      [13..15] opacity  [16] roughness  [17] alpha  [18] eta
 *             (alpha = GGX::roughnessToAlpha(roughness), microfacet.h:44-50,
 *             evaluated once on the host)
 *   light     32 floats  [0] type: 0 area, 1 point, 2 spot, 3 distant,
 *             4 envmap
 *             area/point/spot: [1] triangle index (area)  [2..4] Le or I
 *             [5..7] position  [8] spot cone angle  [9] cosTotalWidth
 *             [10] cosFalloffStart  [11..19] world-to-light 3x3
 *             [20..28] light-to-world 3x3 (spot)  [29..31] spot target
 *             distant: [2..4] L  [5..7] unit direction towards the light
 *             (light/distant.cpp:58-65)  [8..10] from  [11..13] to
 *             envmap:  [1] envmap image index  [2..4] L  [5] width
 *             [6] height  [7] nLightSamples  [8..16] light-to-world 3x3
 *             [17..25] world-to-light 3x3  [26..28] worldUp
 *             Envmaps come after every other light (parser.cpp:45-56).
 *   camera    16 floats  origin(3), camera-to-world 3x3 row-major(9),
 *             tan(vfov/2), aspect, W, H   (camera/pinhole.cpp:8-33)
 */
#ifndef ORION_AMD_H_
#define ORION_AMD_H_
#include <stdint.h>

#define ORION_MAT_FLOATS 24
#ifdef __cplusplus
extern "C" {
#endif

#define ORION_OK 0
#define ORION_ERR_ARG -1
#define ORION_ERR_IO -2
#define ORION_ERR_PARSE -3
#define ORION_ERR_HIP -4
#define ORION_ERR_OOM -5
#define ORION_ERR_UNSUPPORTED -6
#define ORION_ERR_BUSY -7

typedef struct orion_hscene_s *orion_hscene; /* host scene: resolved records */
typedef struct orion_ctx_s *orion_ctx;       /* one GPU + one HIP stream */
typedef struct orion_scene_s *orion_scene;   /* device-resident scene + film */

const char *orion_last_error(void);
const char *orion_version(void);

/* ---------------------------------------------------------------- host scenes */
typedef struct {
  int n_tris, n_mats, n_lights;
  int width, height;
  int integrator;          /* 0 path, 1 sppm (from <Integrator type>) */
  int max_depth;           /* maxDepth */
  int spp;                 /* nSamplesPerPixel (path) */
  int sppm_iterations;     /* numIterations */
  int sppm_photons;        /* photonsPerIteration (-1 -> nPixels) */
  float sppm_radius;       /* initRadius */
} orion_scene_info;

/* XML .ori loader (util/parser.cpp): Camera, Model (mesh | meshHandy),
 * Material (matte | metal | glass | mirror | plastic | uber, named/name),
 * AreaLight, Light (point | spot), Integrator (path | sppm), Sampler (sobol). */
int orion_hscene_load_ori(const char *path, orion_hscene *out);
/* Synthetic scenes: name is one of "cornell", "icosphere", "caustics",
 * "batch_job", "materials", "sky"; params (may be NULL) override defaults:
 *   cornell, materials, sky: [0] width [1] height
 *   icosphere: [0] level [1] width [2] height
 *   caustics:  [0] level [1] width [2] height
 *   batch_job: [0] triangle count [1] width [2] height [3] material mix */
int orion_hscene_generate(const char *name, uint64_t seed, const double *params, int n_params,
                          orion_hscene *out);
int orion_hscene_from_records(int n_tris, const float *tri_pn, const int *tri_mat,
                              const int *tri_light, int n_mats, const float *mats,
                              int n_lights, const float *lights, const float *camera,
                              orion_hscene *out);
int orion_hscene_info(orion_hscene s, orion_scene_info *info);
/* Borrowed pointers into the host scene (valid until destroy). */
int orion_hscene_records(orion_hscene s, const float **tri_pn, const int **tri_mat,
                         const int **tri_light, const float **mats, const float **lights,
                         const float **camera);
/* Writes an .ori (+ .obj for meshes above mesh_inline_limit triangles). */
int orion_hscene_write_ori(orion_hscene s, const char *path, int mesh_inline_limit);
void orion_hscene_destroy(orion_hscene s);
/* No counterpart in orion (diagnostic): builds the BVH that orion_scene_upload
 * would build and checks it structurally on the host -- each node reached once,
 * each triangle in exactly one leaf, every decoded (quantized) child box
 * containing its subtree's vertices.  Fills n_nodes / depth / stack_bound. */
int orion_hscene_bvh_check(orion_hscene s, int64_t *n_nodes, int *depth, int *stack_bound);

/* Resolved camera record from a lookat + fov (transform.cpp:197-217,
 * pinhole.cpp:35-48); exposed for the loader's parity tests. */
int orion_camera_record(const float lookat[9], float fov_deg, int width, int height,
                        float record[16]);
/* createTransform over (key, value) pairs (transform.cpp:218-252): m and mInv, row-major. */
int orion_create_transform(int n, const char *const *keys, const char *const *values,
                           float m[16], float m_inv[16]);
/* Appends an environment light (light/envmap.cpp:9-42, createEnvMap
 * :108-128).  rgb: width x height RGB texels, row 0 first, as loadImage gives
 * them (byte / 255; texture/imagetexture.cpp:22-48); NULL with width =
 * height = 1 is the constant map L.  The map is L x rgb.  light_to_world and
 * world_to_light are row-major 4x4 (their 3x3 parts orient the map). */
int orion_hscene_add_envmap(orion_hscene s, int width, int height, const float *rgb,
                            const float L[3], const float light_to_world[16],
                            const float world_to_light[16]);
/* Texels (before L) of envmap image k: borrowed pointer. */
int orion_hscene_envmap_image(orion_hscene s, int k, int *width, int *height,
                              const float **rgb);
/* Resolved point / spot / distant light record as the loader builds it
 * (light/point.cpp:34-40, light/spot.cpp:46-59, light/distant.cpp:58-65):
 * type 1 point (I, from), 2 spot (I, from, to, angle, delta), 3 distant
 * (L, from, to). */
int orion_light_record(int type, const float I[3], const float from[3], const float to[3],
                       float angle, float delta, float record[32]);

/* ---------------------------------------------------------------- device */
int orion_ctx_create(int device, orion_ctx *out);
int orion_ctx_destroy(orion_ctx c);
int orion_ctx_synchronize(orion_ctx c);
/* Self-check of device arithmetic the traversal relies on (no counterpart in
 * orion): the reciprocal of the triangle and slab tests (the hardware
 * estimate refined by one Newton-Raphson step) against IEEE division, over
 * every float whose magnitude lies in [2^-125, 2^125), on the context's GPU.
 * *mismatches = how many differ in any bit (0 on gfx950; magnitudes outside
 * the range use the division itself). */
int orion_selfcheck_rcp(orion_ctx c, uint64_t *mismatches);

/* Builds the SAH BVH on the host and uploads scene + Sobol tables to HBM.
 * Allocates the film (W*H sum + count) on the device. */
int orion_scene_upload(orion_ctx c, orion_hscene s, orion_scene *out);
int orion_scene_destroy(orion_scene s);
typedef struct {
  int64_t n_nodes, n_tris, node_bytes, tri_bytes, shade_bytes;
  int bvh_depth, node_format;
  double build_ms, upload_ms;
  int stack_bound;          /* traversal stack entries that can be live at once */
  int trace_waves_per_simd; /* occupancy of the closest-hit traversal kernel (default LDS ring) */
} orion_scene_stats;
int orion_scene_get_stats(orion_scene s, orion_scene_stats *st);

/* Batched closest / any hit.  rays: SoA [8][n] = ox oy oz dx dy dz tmin tmax.
 * on_device != 0: all pointers are device pointers (no copies).
 * Outputs: t, prim (-1 = miss), u, v, tie (1 when another triangle was
 * accepted at exactly the winning t); occluded for the any-hit query. */
int orion_intersect_closest(orion_scene s, int n, const float *rays, float *t, int *prim,
                            float *u, float *v, int *tie, int on_device);
int orion_intersect_any(orion_scene s, int n, const float *rays, int *occluded, int on_device);
/* No counterpart in orion: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  orion_intersect_closest with on_device = 1 -- the same
 * results bit for bit, all pointers device pointers, u, v and tie optional
 * -- but it allocates nothing: its traversal runs on a fixed grid whose
 * stack spill area and overflow counter live in one static device array per
 * GPU (32 MiB, held until the process ends).  Limit: scenes whose traversal
 * stack bound (orion_scene_stats.stack_bound, 3 x (BVH depth + 1)) is at
 * most 48, i.e. BVH depth at most 15; larger fail with
 * ORION_ERR_UNSUPPORTED.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (several contexts' streams) race.  Returns after
 * the hits are written (it synchronises the context's stream). */
int orion_intersect_closest_noalloc(orion_scene s, int n, const float *rays, float *t, int *prim, float *u,
                                    float *v, int *tie);

/* Render statistics (device counters). */
typedef struct {
  int64_t paths;          /* camera samples completed */
  int64_t closest_rays;   /* closest-hit casts (camera, bounce, photon) */
  int64_t shadow_rays;    /* any-hit casts (shadow rays, MIS visibility queries) */
  int64_t ref_equiv_rays; /* casts orion itself would make (MIS ray counted twice) */
  int64_t photons;
  int64_t deposits;       /* SPPM photon deposit lookups */
  double ms;              /* device time of the call */
  /* traversal work (for the roofline's algorithmic bytes) */
  int64_t node_visits_closest, tri_tests_closest;
  int64_t node_visits_shadow, tri_tests_shadow;
  /* per-kernel device time, filled when orion_tuning.profile != 0 (HIP
   * events recorded on the stream the kernels run on) */
  double trace_closest_ms, trace_shadow_ms, shade_ms;
  int64_t trace_closest_launches, trace_shadow_launches, shade_launches;
  int64_t stack_overflows;  /* traversal stack entries dropped (0 for a well-formed BVH) */
  /* closest-hit + any-hit queues traced by one launch per bounce (the
   * default; split_trace = 1 launches them separately, timed above) */
  double trace_pair_ms;
  int64_t trace_pair_launches;
  int64_t shaded_segments;  /* path segments shaded (one per live path per bounce) */
  /* SPPM photon pass: photon rays traced (all bounces), their traversal work,
   * in-bounds grid lookups of photon hits after the first, visible-point
   * entries scanned by those lookups; launch times when
   * orion_sppm_debug.profile != 0 (HIP events on the photon pass's stream:
   * traversal, and shading = photon bounces + deposits) */
  int64_t photon_rays, node_visits_photon, tri_tests_photon;
  int64_t photon_lookups, photon_vp_scanned;
  double photon_trace_ms, photon_shade_ms;
  int64_t photon_trace_launches, photon_shade_launches;
  int64_t photon_hits;       /* photon hits after the first bounce, kept for deposits */
  int64_t photon_batch;      /* iterations whose photons are traced together */
} orion_render_stats;

typedef struct {
  int spp_per_wave;     /* plain renders: samples per pixel per wavefront (0 = from wave_paths) */
  int wave_paths;       /* paths per wavefront (0 = 2^26); served waves are capped at 2^25 */
  int rays_per_wave;    /* traversal queue entries per wave64 (0: 512 for one-job waves of 2^25 paths or more, else 256) */
  int use_graph;        /* capture the bounce loop into a hipGraph */
  int profile;          /* time every traversal / shade launch with HIP events */
  int lds_ring;         /* traversal stack entries per lane kept in LDS: 8, 16, 32 (0 = 16) */
  int queue_segment;    /* paths per queue segment, a power of two in [512, 16384] (0 = by wave size) */
  int split_trace;      /* 1: closest-hit and any-hit queues in separate launches (0: one launch) */
  int trace_plan;       /* plain renders' traversal grids: 0 = each bounce's grid sized from the longest
                         * queue segments of the previous call on this scene with the same depth and
                         * queue segment size (a longer segment is still traced in full), 1 = full grids,
                         * 2 = one wave per queue segment at every launch (for tests) */
} orion_tuning;

/* PathTracing::render over samples [k0, k1): accumulates into the scene's film
 * (sum + count per pixel, in sample order per pixel). */
int orion_render_path(orion_scene s, int k0, int k1, int max_depth, const orion_tuning *tune,
                      orion_render_stats *stats);
int orion_film_clear(orion_scene s);
/* rgb_sum: W*H*3 floats, count: W*H ints (host pointers). */
int orion_film_read(orion_scene s, float *rgb_sum, int *count);
/* RenderTarget::getAveraged (texture/rendertarget.cpp:28-37), computed on the
 * device: rgb = sum / count per pixel (W*H*3 floats, host pointer; a pixel
 * with count 0 divides by zero, as in orion). */
int orion_film_read_averaged(orion_scene s, float *rgb);
/* getAveraged of a host (sum, count) film, e.g. a batched or served job's. */
int orion_film_average(int width, int height, const float *rgb_sum, const int *count,
                       float *rgb);
/* Writes an averaged film (W*H*3 floats, row 0 = top):
 *   ORION_IMAGE_BMP  saveAsBmp (texture/texture.cpp:77-137): 32-bit BGRA,
 *                    channels (char)(clamp(v, 0, 1) * 255), 54-byte header
 *   ORION_IMAGE_PFM  FP32 RGB portable float map
 *   ORION_IMAGE_EXR  OpenEXR scanline, uncompressed FLOAT R, G, B */
#define ORION_IMAGE_BMP 0
#define ORION_IMAGE_PFM 1
#define ORION_IMAGE_EXR 2
int orion_film_write(const char *path, int format, int width, int height, const float *rgb);

/* SPPM::render.  hash_size <= 0 -> nPixels (the reference's choice).
 * Optional bit-exact debug outputs (host pointers, may be NULL) for iteration
 * debug_iter: vp_cell_count[hash_size], photon_cell_count[hash_size],
 * M[nPixels], grid_info[8] (gridRes xyz, pair count), grid_bounds[6]. */
typedef struct {
  int debug_iter;
  int *vp_cell_count;
  int *photon_cell_count;
  int *M;
  int *grid_info;
  float *grid_bounds;
  /* The radix-sorted (hash key, pixel) pairs, pairs[2 i] = key, pairs[2 i
   * + 1] = pixel, keys ascending and each key's pixels in the order the
   * reference's per-cell lists receive them (sppm.cpp:170-193); written
   * when there are at most max_pairs (ORION_ERR_ARG otherwise).  cell_range
   * [2 hash_size]: each cell's [start, end) in that order (0, 0 when
   * empty).  Either pointer may be NULL. */
  int *pairs;
  int64_t max_pairs;
  int *cell_range;
  /* != 0: time every photon-pass traversal and shading launch into the
   * render statistics (applies whatever debug_iter is) */
  int profile;
} orion_sppm_debug;
int orion_render_sppm(orion_scene s, int n_iters, int photons_per_iter, float init_radius,
                      int max_depth, int hash_size, const orion_sppm_debug *dbg,
                      orion_render_stats *stats);

/* ---------------------------------------------------------------- batched launch
 * No counterpart in orion: many render jobs -- each its own resident scene,
 * film size, spp and depth -- packed into shared wavefronts: a wave's queue
 * segments each belong to one job, and every bounce's traversal and shading
 * launches serve all of the wave's jobs at once.  A job's camera is its
 * scene's view with the job's film size (aspect = width / height).
 * HBM: a wave of N paths holds 220 B of state per path plus the traversal
 * spill area (60 B per path for the 1.3M-triangle scene: 20 spilled stack
 * entries per traversal lane), so the default 2^26-path wave needs about
 * 19 GB; a wave that does not fit is halved until it does. */
typedef struct {
  orion_scene scene;   /* resident scene (BVH, materials, lights, view) */
  int width, height;   /* film of this job */
  int spp;             /* samples [0, spp) per pixel */
  int max_depth;
} orion_job_desc;

/* Renders every job to completion.  rgb_out[j] (width*height*3 floats) and
 * count_out[j] (width*height ints) receive job j's film sums and counts;
 * each equals what orion_render_path gives for that job alone, bit for bit. */
int orion_render_batch(orion_ctx c, int n_jobs, const orion_job_desc *jobs, float *const *rgb_out,
                       int *const *count_out, const orion_tuning *tune, orion_render_stats *stats);

/* ---------------------------------------------------------------- serving
 * No counterpart in orion: a request loop on its own host thread.  Jobs are
 * admitted when their film fits the HBM budget, merged into the running
 * wavefronts (continuous batching: a job joins at the next wave), and
 * retired when their last sample is accumulated. */
typedef struct orion_server_s *orion_server;
int orion_server_create(orion_ctx c, int64_t hbm_budget_bytes, const orion_tuning *tune, orion_server *out);
/* Wave composition: ORION_SERVE_PACK (default) fills each wave with the jobs
 * of least remaining work first, as many samples each as fit;
 * ORION_SERVE_ARRIVAL fills it the same way in submission order (the oldest
 * job first: no job waits behind later, smaller ones); ORION_SERVE_FAIR
 * gives every admitted job at least one sample per wave (all images refine
 * together).  Whatever the policy a job's film is bit-identical to rendering
 * it alone. */
#define ORION_SERVE_PACK 0
#define ORION_SERVE_FAIR 1
#define ORION_SERVE_ARRIVAL 2
/* ORION_SERVE_AGED: like PACK, but a job's remaining work is credited with
 * the work the device did while it waited (its waiting time x the measured
 * paths/s), so a large job cannot be overtaken indefinitely. */
#define ORION_SERVE_AGED 3
int orion_server_set_policy(orion_server s, int policy);
/* thread-safe; *job_id identifies the request */
int orion_server_submit(orion_server s, const orion_job_desc *job, uint64_t *job_id);
/* done = 1 once finished; then copies the film (either pointer may be NULL)
 * and the request's times in seconds since the server started:
 * times[0] submitted, [1] admitted, [2] finished (its film in host memory:
 * the film of a job that a wave finishes is staged in pinned memory by that
 * wave and copied out once the next wave is enqueued, or at once when no
 * work is left). */
int orion_server_poll(orion_server s, uint64_t job_id, int *done, float *rgb, int *count, double times[3]);
/* Progressive preview of a job: its film (sum + count) after the samples
 * rendered so far; *samples = how many (every pixel's count).  A running job's
 * copy is taken at the next wave boundary (the call waits for it); a queued
 * job returns zeros, a finished one its final film.  The preview equals the
 * film of the same job rendered with spp = *samples. */
int orion_server_peek(orion_server s, uint64_t job_id, float *rgb, int *count, int *samples);
/* blocks until every submitted job has finished */
int orion_server_drain(orion_server s);
typedef struct {
  int64_t jobs_done, waves, paths, closest_rays, shadow_rays;
  int active_jobs, queued_jobs;
  int64_t hbm_in_use;
  double busy_ms;
  int64_t mixed_waves;  /* waves that carried more than one job */
  int max_jobs_in_wave;
  int64_t hbm_peak;      /* largest film memory reserved at once (admission) */
} orion_server_stats;
int orion_server_get_stats(orion_server s, orion_server_stats *st);
int orion_server_destroy(orion_server s);

/* ---------------------------------------------------------------- autotuner
 * No counterpart in orion.  Times short batched renders of `jobs` (each
 * cut to its first two samples per pixel) by coordinate descent over the
 * wave size, traversal chunk (queue entries per wave64), LDS stack ring,
 * queue segment size, graph capture and split traversal launches, and
 * returns the fastest setting for orion_render_batch / orion_server_create;
 * with one job whose film is its scene's, also for orion_render_path. */
int orion_autotune(orion_ctx c, int n_jobs, const orion_job_desc *jobs, orion_tuning *best);

#ifdef __cplusplus
}
#endif
#endif
