// Host scene: the resolved records a render consumes (see include/orion_amd.h
// for the record layouts), plus the integrator settings read from the .ori.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/orion_amd.h"

namespace orion_amd {

enum { MAT_MATTE = 0, MAT_METAL = 1, MAT_GLASS = 2, MAT_MIRROR = 3, MAT_PLASTIC = 4, MAT_UBER = 5 };
enum { LIGHT_AREA = 0, LIGHT_POINT = 1, LIGHT_SPOT = 2 };
constexpr int kTriFloats = 18, kMatFloats = ORION_MAT_FLOATS, kLightFloats = 32, kCamFloats = 16;

struct HostScene {
  std::vector<float> tri_pn;    // n x 18
  std::vector<int> tri_mat;     // n
  std::vector<int> tri_light;   // n (area light index or -1)
  std::vector<float> mats;      // m x kMatFloats
  std::vector<float> lights;    // l x 32
  float camera[kCamFloats] = {0};
  orion_scene_info info{};
  // Writer support: keeps the lookat/fov the camera was built from.
  float lookat[9] = {0, 0, 1, 0, 0, 0, 0, 1, 0};
  float fov = 40.f;
  // Per-material names for the writer (optional).
  int n_tris() const { return (int)tri_mat.size(); }
  int n_mats() const { return (int)(mats.size() / kMatFloats); }
  int n_lights() const { return (int)(lights.size() / kLightFloats); }
  void refresh_info();
};

// Record builders shared by the loader and the generators.
void camera_record(const float lookat[9], float fov_deg, int W, int H, float rec[16]);
float roughness_to_alpha(float roughness);
void matte_record(float r, float g, float b, float rec[kMatFloats], float sigma = 0.f);
void metal_record(const float eta[3], const float k[3], float roughness, float rec[kMatFloats]);
void mirror_record(const float Kr[3], float rec[kMatFloats]);
void plastic_record(const float Kd[3], const float Ks[3], float roughness, float rec[kMatFloats]);
void uber_record(const float Kd[3], const float Ks[3], const float Kr[3], const float Kt[3],
                 const float opacity[3], float roughness, float eta, float rec[kMatFloats]);
bool metal_preset(const std::string &name, float eta[3], float k[3]);
void glass_record(const float Kr[3], const float Kt[3], float roughness, float eta, float rec[kMatFloats]);
void point_light_record(const float I[3], const float from[3], float rec[32]);
void spot_light_record(const float I[3], const float from[3], const float to[3], float angle,
                       float delta, float rec[32]);
void area_light_record(int tri, const float Le[3], float rec[32]);

// .ori loading / writing (ori_loader.cpp)
int load_ori(const std::string &path, HostScene *out, std::string *err);
int write_ori(const HostScene &s, const std::string &path, int inline_limit, std::string *err);

// Synthetic scenes (scenegen.cpp)
int generate_scene(const std::string &name, uint64_t seed, const double *params, int n_params,
                   HostScene *out, std::string *err);

void set_error(const std::string &msg);

}  // namespace orion_amd
