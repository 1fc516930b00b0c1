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

// distant.cpp:58-65 + distant.h:23-24: dir = normalize(to - from), then the
// light-to-world transform translate(from) is applied to it (a vector: the
// translation drops out) and the result normalised again.
void distant_light_record(const float L[3], const float from[3], const float to[3], float rec[32]) {
  std::memset(rec, 0, 32 * sizeof(float));
  rec[0] = LIGHT_DISTANT;
  rec[1] = -1;
  F3 pf = f3(from[0], from[1], from[2]), pt = f3(to[0], to[1], to[2]);
  F3 d = normalize(sub(pt, pf));
  d = normalize(apply_vector(translate(pf).m, d));
  rec[2] = L[0]; rec[3] = L[1]; rec[4] = L[2];
  rec[5] = d.x; rec[6] = d.y; rec[7] = d.z;
  for (int i = 0; i < 3; ++i) { rec[8 + i] = from[i]; rec[11 + i] = to[i]; }
}

// createEnvMap (envmap.cpp:108-128): t = rotate(degrees(acos(worldUp . z)),
// z x worldUp) unless worldUp is within acos(0.99) of +-z (a missing worldUp
// normalises to NaN and fails both tests: no rotation); light-to-world =
// t * transform.
void envmap_orientation(const float xm[16], const float xmi[16], const float world_up[3], float l2w[16],
                        float w2l[16]) {
  Xform x;
  std::memcpy(x.m.m, xm, 16 * sizeof(float));
  std::memcpy(x.mInv.m, xmi, 16 * sizeof(float));
  F3 up = normalize(f3(world_up[0], world_up[1], world_up[2]));
  F3 ez = f3(0, 0, 1);
  float cosAngle = dot(up, ez);
  Xform t;
  if ((double)cosAngle < 0.99 && (double)cosAngle > -0.99) {
    float angle = std::acos(dot(up, ez));
    F3 axis = cross(ez, up);
    t = rotate(angle * static_cast<float>(57.295779513082320876798154814105), axis);  // degrees()
  }
  Xform r = t * x;
  std::memcpy(l2w, r.m.m, 16 * sizeof(float));
  std::memcpy(w2l, r.mInv.m, 16 * sizeof(float));
}

void envmap_light_record(int image, const float L[3], int width, int height, int n_samples, const float l2w[16],
                         const float w2l[16], const float world_up[3], float rec[32]) {
  std::memset(rec, 0, 32 * sizeof(float));
  rec[0] = LIGHT_ENVMAP;
  rec[1] = (float)image;
  rec[2] = L[0]; rec[3] = L[1]; rec[4] = L[2];
  rec[5] = (float)width;
  rec[6] = (float)height;
  rec[7] = (float)(n_samples > 1 ? n_samples : 1);  // Light::nSamples = max(1, n)
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      rec[8 + 3 * r + c] = l2w[4 * r + c];
      rec[17 + 3 * r + c] = w2l[4 * r + c];
    }
  for (int i = 0; i < 3; ++i) rec[26 + i] = world_up ? world_up[i] : 0.f;
}

namespace {
// Distribution1D (sampling.h:76-90) appended as [func][cdf][funcInt]
void append_distribution(const float *f, int n, std::vector<float> *o) {
  size_t base = o->size();
  o->insert(o->end(), f, f + n);
  std::vector<float> cdf(n + 1);
  cdf[0] = 0;
  for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / n;
  float funcInt = cdf[n];
  if (funcInt == 0) {
    for (int i = 1; i < n + 1; ++i) cdf[i] = float(i) / float(n);
  } else {
    for (int i = 1; i < n + 1; ++i) cdf[i] /= funcInt;
  }
  o->insert(o->end(), cdf.begin(), cdf.end());
  o->push_back(funcInt);
  (void)base;
}
float intensity3(const float *s) { return 0.212671f * s[0] + 0.715160f * s[1] + 0.072169f * s[2]; }
}  // namespace

bool build_env_tables(const HostScene &s, EnvTables *t, std::string *err) {
  *t = EnvTables();
  for (int l = 0; l < s.n_lights(); ++l) {
    const float *r = &s.lights[32 * (size_t)l];
    if ((int)r[0] != LIGHT_ENVMAP) continue;
    int k = (int)r[1];
    if (k < 0 || k >= (int)s.env.size()) { *err = "envmap light without an image"; return false; }
    const EnvImage &img = s.env[k];
    int w = img.width, h = img.height;
    if (w <= 0 || h <= 0 || img.rgb.size() != 3 * (size_t)w * h || (int)r[5] != w || (int)r[6] != h) {
      *err = "envmap image size does not match its light record";
      return false;
    }
    size_t toff = t->tex.size(), doff = t->dist.size();
    t->meta.insert(t->meta.end(), {l, w, h, (int)toff, (int)doff});
    if (t->tex.size() + 3 * (size_t)w * h > (size_t)INT32_MAX) { *err = "envmap texels exceed 2^31"; return false; }
    // envmap.cpp:14-20: data[i] *= L
    for (size_t i = 0; i < (size_t)w * h; ++i)
      for (int c = 0; c < 3; ++c) t->tex.push_back(img.rgb[3 * i + c] * r[2 + c]);
    const float *tx = &t->tex[toff];
    // envmap.cpp:27-39: img = intensity * sin(pi v), aveSpectrum = mean texel
    std::vector<float> f((size_t)w * h);
    float ave[3] = {0, 0, 0};
    for (int j = 0; j < h; ++j) {
      float v = (j + 0.5f) / (float)h;
      float sinTheta = std::sin(3.14159265358979323846f * v);
      for (int i = 0; i < w; ++i) {
        const float *sx = tx + 3 * ((size_t)j * w + i);
        f[i + (size_t)j * w] = intensity3(sx) * sinTheta;
        for (int c = 0; c < 3; ++c) ave[c] += sx[c];
      }
    }
    for (int c = 0; c < 3; ++c) t->ave.push_back(ave[c] / (w * h));
    // Distribution2D (sampling.cpp:63-76): marginal over the rows' funcInt
    std::vector<float> rows;
    std::vector<float> marg(h);
    for (int v = 0; v < h; ++v) {
      size_t at = rows.size();
      append_distribution(&f[(size_t)v * w], w, &rows);
      marg[v] = rows[at + 2 * (size_t)w + 1];
    }
    append_distribution(marg.data(), h, &t->dist);
    t->dist.insert(t->dist.end(), rows.begin(), rows.end());
    if (t->dist.size() > (size_t)INT32_MAX) { *err = "envmap tables exceed 2^31"; return false; }
    (void)doff;
  }
  return true;
}

void world_sphere(const HostScene &s, float center[3], float *radius) {
  auto mn = [](float a, float b) { return (b < a) ? b : a; };  // std::min
  auto mx = [](float a, float b) { return (a < b) ? b : a; };  // std::max
  float lo[3], hi[3];
  for (int c = 0; c < 3; ++c) { lo[c] = 3.40282347e+38f; hi[c] = -3.40282347e+38f; }
  for (int i = 0; i < s.n_tris(); ++i) {
    const float *q = &s.tri_pn[18 * (size_t)i];
    for (int c = 0; c < 3; ++c) {
      // Union(Bounds3f(p0, p1), p2) (triangle.cpp:60-66), then Union(bound, that)
      float tlo = mn(mn(q[c], q[3 + c]), q[6 + c]);
      float thi = mx(mx(q[c], q[3 + c]), q[6 + c]);
      lo[c] = mn(lo[c], tlo);
      hi[c] = mx(hi[c], thi);
    }
  }
  float inv = (float)1 / 2;
  bool inside = true;
  float d2 = 0;
  for (int c = 0; c < 3; ++c) {
    center[c] = inv * (lo[c] + hi[c]);
    inside = inside && center[c] >= lo[c] && center[c] <= hi[c];
  }
  float dx = center[0] - hi[0], dy = center[1] - hi[1], dz = center[2] - hi[2];
  d2 = dx * dx + dy * dy + dz * dz;
  *radius = inside ? std::sqrt(d2) : 0.f;
}

void light_powers(const HostScene &s, const EnvTables &env, float radius, std::vector<float> *rgb) {
  const float pi = 3.14159265358979323846f;
  rgb->assign(3 * (size_t)s.n_lights(), 0.f);
  int e = 0;
  for (int l = 0; l < s.n_lights(); ++l) {
    const float *L = &s.lights[32 * (size_t)l];
    float *o = &(*rgb)[3 * (size_t)l];
    int type = (int)L[0];
    if (type == LIGHT_AREA) {  // DiffuseAreaLight::power = Le * area * pi (diffuse.cpp:8-11)
      int t = (int)L[1];
      const float *q = &s.tri_pn[18 * (size_t)t];
      float e1x = q[3] - q[0], e1y = q[4] - q[1], e1z = q[5] - q[2];
      float e2x = q[6] - q[0], e2y = q[7] - q[1], e2z = q[8] - q[2];
      float cx = (e1y * e2z) - (e1z * e2y), cy = (e1z * e2x) - (e1x * e2z), cz = (e1x * e2y) - (e1y * e2x);
      float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
      for (int c = 0; c < 3; ++c) o[c] = L[2 + c] * area * pi;
    } else if (type == LIGHT_POINT) {  // PointLight::power = 4 pi I (point.cpp:6-9)
      const float k = 4 * pi;
      for (int c = 0; c < 3; ++c) o[c] = L[2 + c] * k;
    } else if (type == LIGHT_SPOT) {  // spot.cpp:6-9
      float f = (1 - .5f * (L[10] + L[9]));
      for (int c = 0; c < 3; ++c) o[c] = L[2 + c] * 2 * pi * f;
    } else if (type == LIGHT_DISTANT) {  // DistantLight::power = L pi r^2 (distant.cpp:11-14)
      for (int c = 0; c < 3; ++c) o[c] = L[2 + c] * pi * radius * radius;
    } else {  // Envmap::power = pi r^2 aveSpectrum (envmap.cpp:51-54)
      float k = pi * radius * radius;
      for (int c = 0; c < 3; ++c) o[c] = env.ave[3 * (size_t)e + c] * k;
      ++e;
    }
  }
}

}  // namespace orion_amd
