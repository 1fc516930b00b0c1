// C ABI for host scenes: load / generate / inspect / write.
#include <cstring>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"
#include "bvh.h"
#include "host_math.h"
#include "host_scene.h"

namespace orion_amd {
const char *last_error_cstr();
bool create_transform(const std::vector<std::pair<std::string, std::string>> &pv, Xform *out, std::string *err);
}
using namespace orion_amd;

extern "C" {

const char *orion_last_error(void) { return last_error_cstr(); }
const char *orion_version(void) { return "orion_amd 0.1 (gfx950)"; }

int orion_hscene_load_ori(const char *path, orion_hscene *out) {
  if (!path || !out) { set_error("null argument"); return ORION_ERR_ARG; }
  auto *s = new HostScene();
  std::string err;
  int rc = load_ori(path, s, &err);
  if (rc) { delete s; set_error(err); return rc; }
  *out = reinterpret_cast<orion_hscene>(s);
  return ORION_OK;
}

int orion_hscene_generate(const char *name, uint64_t seed, const double *params, int n_params, orion_hscene *out) {
  if (!name || !out) { set_error("null argument"); return ORION_ERR_ARG; }
  auto *s = new HostScene();
  std::string err;
  int rc = generate_scene(name, seed, params, n_params, s, &err);
  if (rc) { delete s; set_error(err); return rc; }
  *out = reinterpret_cast<orion_hscene>(s);
  return ORION_OK;
}

int orion_hscene_from_records(int n_tris, const float *tri_pn, const int *tri_mat, const int *tri_light, int n_mats,
                              const float *mats, int n_lights, const float *lights, const float *camera,
                              orion_hscene *out) {
  if (n_tris < 0 || n_mats < 0 || n_lights < 0 || !camera || !out || (n_tris && (!tri_pn || !tri_mat || !tri_light)) ||
      (n_mats && !mats) || (n_lights && !lights)) {
    set_error("bad argument");
    return ORION_ERR_ARG;
  }
  auto *s = new HostScene();
  s->tri_pn.assign(tri_pn, tri_pn + 18 * (size_t)n_tris);
  s->tri_mat.assign(tri_mat, tri_mat + n_tris);
  s->tri_light.assign(tri_light, tri_light + n_tris);
  s->mats.assign(mats, mats + kMatFloats * (size_t)n_mats);
  s->lights.assign(lights, lights + 32 * (size_t)n_lights);
  std::memcpy(s->camera, camera, 16 * sizeof(float));
  s->refresh_info();
  *out = reinterpret_cast<orion_hscene>(s);
  return ORION_OK;
}

int orion_hscene_info(orion_hscene h, orion_scene_info *info) {
  if (!h || !info) { set_error("null argument"); return ORION_ERR_ARG; }
  HostScene *s = reinterpret_cast<HostScene *>(h);
  s->refresh_info();
  *info = s->info;
  return ORION_OK;
}

int orion_hscene_records(orion_hscene h, const float **tri_pn, const int **tri_mat, const int **tri_light,
                         const float **mats, const float **lights, const float **camera) {
  if (!h) { set_error("null scene"); return ORION_ERR_ARG; }
  HostScene *s = reinterpret_cast<HostScene *>(h);
  if (tri_pn) *tri_pn = s->tri_pn.data();
  if (tri_mat) *tri_mat = s->tri_mat.data();
  if (tri_light) *tri_light = s->tri_light.data();
  if (mats) *mats = s->mats.data();
  if (lights) *lights = s->lights.data();
  if (camera) *camera = s->camera;
  return ORION_OK;
}

int orion_hscene_write_ori(orion_hscene h, const char *path, int inline_limit) {
  if (!h || !path) { set_error("null argument"); return ORION_ERR_ARG; }
  std::string err;
  int rc = write_ori(*reinterpret_cast<HostScene *>(h), path, inline_limit, &err);
  if (rc) set_error(err);
  return rc;
}

void orion_hscene_destroy(orion_hscene h) { delete reinterpret_cast<HostScene *>(h); }

int orion_hscene_bvh_check(orion_hscene h, int64_t *n_nodes, int *depth, int *stack_bound) {
  if (!h) { set_error("null scene"); return ORION_ERR_ARG; }
  const HostScene &s = *reinterpret_cast<HostScene *>(h);
  int n = (int)s.tri_mat.size();
  Bvh b;
  build_bvh(s.tri_pn.data(), n, &b);
  std::string err = check_bvh(b, s.tri_pn.data(), n);
  if (n_nodes) *n_nodes = (int64_t)b.nodes.size();
  if (depth) *depth = b.depth;
  if (stack_bound) *stack_bound = b.stack_bound;
  if (!err.empty()) { set_error("BVH check: " + err); return ORION_ERR_ARG; }
  return ORION_OK;
}

int orion_camera_record(const float lookat[9], float fov, int W, int H, float rec[16]) {
  if (!lookat || !rec || W <= 0 || H <= 0) { set_error("bad argument"); return ORION_ERR_ARG; }
  camera_record(lookat, fov, W, H, rec);
  return ORION_OK;
}

int orion_create_transform(int n, const char *const *keys, const char *const *values, float m[16], float mi[16]) {
  if (n < 0 || (n && (!keys || !values)) || !m || !mi) { set_error("bad argument"); return ORION_ERR_ARG; }
  std::vector<std::pair<std::string, std::string>> pv;
  for (int i = 0; i < n; ++i) pv.emplace_back(keys[i], values[i]);
  Xform t;
  std::string err;
  if (!create_transform(pv, &t, &err)) { set_error(err); return ORION_ERR_PARSE; }
  std::memcpy(m, t.m.m, 16 * sizeof(float));
  std::memcpy(mi, t.mInv.m, 16 * sizeof(float));
  return ORION_OK;
}

int orion_light_record(int type, const float I[3], const float from[3], const float to[3], float angle, float delta,
                       float rec[32]) {
  if (!I || !from || !rec) { set_error("bad argument"); return ORION_ERR_ARG; }
  if (type == LIGHT_POINT) point_light_record(I, from, rec);
  else if (type == LIGHT_SPOT || type == LIGHT_DISTANT) {
    if (!to) { set_error("spot and distant lights need to"); return ORION_ERR_ARG; }
    if (type == LIGHT_SPOT) spot_light_record(I, from, to, angle, delta, rec);
    else distant_light_record(I, from, to, rec);
  } else {
    set_error("type must be 1 (point), 2 (spot) or 3 (distant)");
    return ORION_ERR_ARG;
  }
  return ORION_OK;
}

int orion_hscene_add_envmap(orion_hscene h, int width, int height, const float *rgb, const float L[3],
                            const float light_to_world[16], const float world_to_light[16]) {
  if (!h || !L || !light_to_world || !world_to_light || width <= 0 || height <= 0 ||
      (!rgb && (width != 1 || height != 1)) || (int64_t)width * height > (1 << 26)) {
    set_error("bad argument");
    return ORION_ERR_ARG;
  }
  HostScene *s = reinterpret_cast<HostScene *>(h);
  EnvImage img;
  img.width = width;
  img.height = height;
  if (rgb) img.rgb.assign(rgb, rgb + 3 * (size_t)width * height);
  else img.rgb.assign(3, 1.f);
  img.has_xform_params = false;
  float rec[32];
  envmap_light_record((int)s->env.size(), L, width, height, 1, light_to_world, world_to_light, nullptr, rec);
  s->lights.insert(s->lights.end(), rec, rec + 32);
  s->env.push_back(std::move(img));
  s->refresh_info();
  return ORION_OK;
}

int orion_hscene_envmap_image(orion_hscene h, int k, int *width, int *height, const float **rgb) {
  if (!h) { set_error("null scene"); return ORION_ERR_ARG; }
  HostScene *s = reinterpret_cast<HostScene *>(h);
  if (k < 0 || k >= (int)s->env.size()) { set_error("envmap image index out of range"); return ORION_ERR_ARG; }
  if (width) *width = s->env[k].width;
  if (height) *height = s->env[k].height;
  if (rgb) *rgb = s->env[k].rgb.data();
  return ORION_OK;
}

}  // extern "C"
