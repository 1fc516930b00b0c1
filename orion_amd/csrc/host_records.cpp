// Record builders: resolve orion's scene objects (camera, materials, lights)
// into the flat records the renderer consumes.  Host-only, runs once per scene.
#include <cmath>
#include <cstring>
#include <string>

#include "host_math.h"
#include "host_scene.h"
#include "metal_presets.inc"

namespace orion_amd {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error_cstr() { return g_last_error.c_str(); }

void HostScene::refresh_info() {
  info.n_tris = n_tris();
  info.n_mats = n_mats();
  info.n_lights = n_lights();
  info.width = (int)camera[14];
  info.height = (int)camera[15];
}

// Pinhole camera: camera-to-world = inverse(lookAt) (transform.cpp:197-217),
// origin = that transform applied to (0,0,0) (pinhole.cpp:31), tan_half_fov =
// tan(radians(vfov * 0.5)) (pinhole.h:25), aspect = W / H (camera.h:31).
void camera_record(const float la[9], float fov, int W, int H, float rec[16]) {
  Xform t = look_at(f3(la[0], la[1], la[2]), f3(la[3], la[4], la[5]), f3(la[6], la[7], la[8]));
  // createTransform starts from the identity: t = lookAt * identity.
  t = t * Xform();
  F3 o = apply_point(t.m, f3(0, 0, 0));
  rec[0] = o.x; rec[1] = o.y; rec[2] = o.z;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) rec[3 + 3 * r + c] = t.m.m[r][c];
  rec[12] = std::tan(radians(fov * 0.5f));
  rec[13] = (float)W / H;
  rec[14] = (float)W;
  rec[15] = (float)H;
}

// microfacet.h:44-50
float roughness_to_alpha(float roughness) {
  roughness = roughness < (float)1e-3 ? (float)1e-3 : roughness;
  float x = std::log(roughness);
  return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x +
         0.000640711f * x * x * x * x;
}

void matte_record(float r, float g, float b, float rec[kMatFloats], float sigma) {
  std::memset(rec, 0, kMatFloats * sizeof(float));
  rec[0] = MAT_MATTE;
  rec[1] = r; rec[2] = g; rec[3] = b;
  rec[4] = sigma;
}

void metal_record(const float eta[3], const float k[3], float roughness, float rec[kMatFloats]) {
  std::memset(rec, 0, kMatFloats * sizeof(float));
  rec[0] = MAT_METAL;
  for (int i = 0; i < 3; ++i) { rec[1 + i] = eta[i]; rec[4 + i] = k[i]; }
  rec[7] = roughness;
  rec[8] = roughness_to_alpha(roughness);
}

bool metal_preset(const std::string &name, float eta[3], float k[3]) {
  for (const auto &p : orion_metal_presets) {
    if (name == p.name) {
      std::memcpy(eta, p.eta, 3 * sizeof(float));
      std::memcpy(k, p.k, 3 * sizeof(float));
      return true;
    }
  }
  return false;
}

void glass_record(const float Kr[3], const float Kt[3], float roughness, float eta, float rec[kMatFloats]) {
  std::memset(rec, 0, kMatFloats * sizeof(float));
  rec[0] = MAT_GLASS;
  for (int i = 0; i < 3; ++i) { rec[1 + i] = Kr[i]; rec[4 + i] = Kt[i]; }
  rec[7] = roughness;
  rec[8] = roughness_to_alpha(roughness);
  rec[9] = eta;
}

void mirror_record(const float Kr[3], float rec[kMatFloats]) {  // material/mirror.cpp:7-14
  std::memset(rec, 0, kMatFloats * sizeof(float));
  rec[0] = MAT_MIRROR;
  for (int i = 0; i < 3; ++i) rec[1 + i] = Kr[i];
}

void plastic_record(const float Kd[3], const float Ks[3], float roughness, float rec[kMatFloats]) {
  std::memset(rec, 0, kMatFloats * sizeof(float));  // material/plastic.cpp:8-24
  rec[0] = MAT_PLASTIC;
  for (int i = 0; i < 3; ++i) { rec[1 + i] = Kd[i]; rec[4 + i] = Ks[i]; }
  rec[7] = roughness;
  rec[8] = roughness_to_alpha(roughness);
}

void uber_record(const float Kd[3], const float Ks[3], const float Kr[3], const float Kt[3],
                 const float opacity[3], float roughness, float eta, float rec[kMatFloats]) {
  std::memset(rec, 0, kMatFloats * sizeof(float));  // material/uber.cpp:9-42
  rec[0] = MAT_UBER;
  for (int i = 0; i < 3; ++i) {
    rec[1 + i] = Kd[i]; rec[4 + i] = Ks[i]; rec[7 + i] = Kr[i]; rec[10 + i] = Kt[i]; rec[13 + i] = opacity[i];
  }
  rec[16] = roughness;
  rec[17] = roughness_to_alpha(roughness);
  rec[18] = eta;
}

// point.cpp:34-40: light-to-world = translate(from); p = that applied to 0.
void point_light_record(const float I[3], const float from[3], float rec[32]) {
  std::memset(rec, 0, 32 * sizeof(float));
  rec[0] = LIGHT_POINT;
  rec[1] = -1;
  Xform l2w = translate(f3(from[0], from[1], from[2]));
  F3 p = apply_point(l2w.m, f3(0, 0, 0));
  rec[2] = I[0]; rec[3] = I[1]; rec[4] = I[2];
  rec[5] = p.x; rec[6] = p.y; rec[7] = p.z;
}

// spot.cpp:46-59 + spot.h:22-25.  cosFalloffStart is initialised from itself
// in the reference (undefined); this build pins it to cos(radians(0)) = 1,
// i.e. the falloff spans the whole cone -- the value the reference produces
// when the object's storage starts zeroed (checked against oracle/_ref).
void spot_light_record(const float I[3], const float from[3], const float to[3], float angle,
                       float delta, float rec[32]) {
  (void)delta;
  std::memset(rec, 0, 32 * sizeof(float));
  rec[0] = LIGHT_SPOT;
  rec[1] = -1;
  F3 pf = f3(from[0], from[1], from[2]), pt = f3(to[0], to[1], to[2]);
  F3 dir = normalize(sub(pt, pf));
  F3 x, y;
  coordinate_system(dir, &x, &y);
  Xform dirToZ(Mat4::make(x.x, x.y, x.z, 0.0f, y.x, y.y, y.z, 0.0f, dir.x, dir.y, dir.z, 0.0f, 0.0f,
                          0.0f, 0.0f, 1.0f));
  Xform l2w = translate(pf) * inverse(dirToZ);
  Xform w2l = inverse(l2w);
  F3 p = apply_point(l2w.m, f3(0, 0, 0));
  rec[2] = I[0]; rec[3] = I[1]; rec[4] = I[2];
  rec[5] = p.x; rec[6] = p.y; rec[7] = p.z;
  rec[8] = angle;  // kept for the .ori writer (round trip)
  rec[9] = std::cos(radians(angle));
  rec[10] = std::cos(radians(0.0f));
  rec[29] = to[0]; rec[30] = to[1]; rec[31] = to[2];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      rec[11 + 3 * r + c] = w2l.m.m[r][c];
      rec[20 + 3 * r + c] = l2w.m.m[r][c];
    }
}

void area_light_record(int tri, const float Le[3], float rec[32]) {
  std::memset(rec, 0, 32 * sizeof(float));
  rec[0] = LIGHT_AREA;
  rec[1] = (float)tri;
  rec[2] = Le[0]; rec[3] = Le[1]; rec[4] = Le[2];
}

}  // namespace orion_amd
