/* Oracle: CPU restatement of orion's path-tracing and SPPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (orion_amd/) links,
 * includes or calls this library; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, as the checker and the CPU baseline.
 *
 * The renderer core takes a *resolved* scene (world-space triangles, resolved
 * material and light records, camera record) in the record layout documented
 * in include/orion_amd.h, so that the GPU path and this oracle consume
 * byte-identical inputs.  Every function below cites the reference file:line
 * whose behaviour it restates (paths relative to the reference's src/).
 *
 * Parity pins (see the header of oracle.cpp for the per-function table):
 * Sobol, camera rays, BSDF f/sample/pdf, Fresnel, point/spot lights and
 * uniformSampleTriangle are checked against the reference's own compiled
 * sources (oracle/_ref); closest-hit is checked by brute force; the
 * integrators by closed forms (furnace, constant-emission) and invariants.
 */
#ifndef ORION_ORACLE_H_
#define ORION_ORACLE_H_
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Scene handle.  tri_pn: n_tris x 18 floats (p0,p1,p2,n0,n1,n2, world space).
 * tri_mat: material index per triangle.  tri_light: area-light index or -1.
 * mats: n_mats x 24 floats.  lights: n_lights x 32 floats.  camera: 16 floats.
 * Envmap images: image_wh[2k], image_wh[2k+1] = width, height of image k,
 * image_rgb[k] its texels (before L).  Returns NULL on a bad image index. */
void *oracle_scene_create(int n_tris, const float *tri_pn, const int *tri_mat,
                          const int *tri_light, int n_mats, const float *mats,
                          int n_lights, const float *lights, const float *camera,
                          int n_images, const int *image_wh, const float *const *image_rgb);
void oracle_scene_destroy(void *scene);

/* Closest / any hit over a batch of rays (rays: n x 8 = o, d, tmin, tmax).
 * mode 0 = closest (lexicographic min of (t, prim)), 1 = any.
 * brute != 0 tests every triangle instead of walking the oracle BVH.
 * ties[i] = 1 when another triangle was accepted at exactly the winning t. */
int oracle_intersect(void *scene, int n, const float *rays, int mode, int brute,
                     float *t, int *prim, float *u, float *v, int *ties);

/* PathTracing::render (integrator/path.cpp:15-37) over sample range
 * [k0, k1) and rows [row0, row1).  rgb_sum (W*H*3) and count (W*H) are
 * accumulated in place (caller zeroes them).  threads <= 0 = all cores.
 * counters (may be NULL): [0] closest casts, [1] shadow casts, [2] paths. */
int oracle_render_path(void *scene, int k0, int k1, int row0, int row1,
                       int max_depth, float *rgb_sum, int *count, int threads,
                       int64_t *counters);

/* Per-sample radiance for a list of (pixel, sample k) pairs; used by tests. */
int oracle_path_samples(void *scene, int n, const int *pixel, const int *k,
                        int max_depth, float *L);

/* SPPM::render (integrator/sppm.cpp:58-322).  hash_size <= 0 means the
 * reference's nPixels.  Debug outputs (any may be NULL) are taken at
 * iteration debug_iter: vp_cell_count[hash_size] (visible-point entries per
 * hash cell), photon_cell_count[hash_size] (in-bounds photon deposits per
 * cell), M_debug[nPixels] (photon count per pixel before the update),
 * grid_info[8] = gridRes xyz, n_pairs; grid_bounds[6] = bounds lo, hi;
 * vp_out[nPixels * 5] = visible point x, y, z, radius, valid;
 * pairs_out[2 * n_pairs] = the per-cell lists (sppm.cpp:170-193) as (hash
 * key, pixel) pairs, keys ascending, each cell's pixels in insertion order
 * (pixel order: the lists are built serially), written when n_pairs <=
 * max_pairs. */
int oracle_render_sppm(void *scene, int n_iters, int photons_per_iter,
                       float init_radius, int max_depth, int hash_size,
                       float *rgb_sum, int *count, int threads, int debug_iter,
                       int *vp_cell_count, int *photon_cell_count, int *M_debug,
                       int *grid_info, float *grid_bounds, float *vp_out,
                       int *pairs_out, int64_t max_pairs);

/* Building blocks exposed for the unit pins. */
float oracle_sobol(uint64_t index, int dim);
/* Camera ray for pixel (i, j) with Sobol index `index` (pinhole.cpp:8-33).
 * out: o(3), d(3). */
void oracle_camera_ray(void *scene, int i, int j, uint64_t index, float *out);
/* BSDF of material m at a shading frame.  frame: ng(3) ns(3) u(3) v(3).
 * mode 0 = Path, 1 = Photon.  eval: out f(3), pdf.  sample: in u(2),
 * out f(3), wi(3), pdf, type. */
void oracle_bsdf_f_pdf(void *scene, int m, int mode, const float *frame,
                       const float *wi, const float *wo, int type_mask,
                       float *f, float *pdf);
void oracle_bsdf_sample(void *scene, int m, int mode, const float *frame,
                        const float *wo, const float *u, int type_mask,
                        float *f, float *wi, float *pdf, int *sampled_type);
/* Light::sample_Li at a reference point p with normal ng.
 * out: Li(3), wi(3), pdf, shadow ray o(3), d(3), tmax. */
void oracle_light_sample_li(void *scene, int l, const float *p, const float *ng,
                            const float *u, float *out);
/* Light::sample_Le.  out: Le(3), o(3), d(3), nLight(3), pdfPos, pdfDir. */
void oracle_light_sample_le(void *scene, int l, const float *u1, const float *u2,
                            float *out);
/* Envmap light l: Le of a ray direction and pdf_Li of a direction
 * (light/envmap.cpp:44-82); any light's power (Light::power); the world
 * bounding sphere (center xyz, radius). */
void oracle_env_le(void *scene, int l, const float *d, float *out);
float oracle_env_pdf_li(void *scene, int l, const float *wi);
void oracle_light_power(void *scene, int l, float *out);
void oracle_world_sphere(void *scene, float *out);
/* sin/cos as the renderer uses them. */
float oracle_sin(float x);
float oracle_cos(float x);
/* Warps of sampler/sampling.cpp for u = (u[0], u[1]); out (13 floats, the
 * layout of the _ref harness's ref_warps): uniformSampleTriangle b0 b1,
 * uniformSampleDisk x y, cosineSampleHemisphere xyz, uniformSampleSphere xyz,
 * uniformSampleCone(cos_theta_max) xyz. */
void oracle_warps(const float *u, float cos_theta_max, float *out);
/* offsetRayOrigin(p, ng, d) (core/geometry.h:850-863): the origin of
 * Intersection::spawnRay (core/intersection.h:40-49). */
void oracle_spawn(const float *p, const float *ng, const float *d, float *o3);
/* Distribution1D over func[0..n) (sampler/sampling.h:71-125) at u; out:
 * [0] sampleDiscrete offset, [1] its pdf, [2] sampleContinuous value,
 * [3] its pdf, [4] its offset. */
void oracle_distribution1d(int n, const float *func, float u, float *out);

#ifdef __cplusplus
}
#endif
#endif
