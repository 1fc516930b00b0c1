// Synthetic scenes for the path's workloads (all seeded, all deterministic):
//   cornell    5 walls of the unit cube + 2 rotated boxes + a 0.25 x 0.25
//              ceiling light (Le 15); 256^2, 16 spp, depth 5
//   icosphere  level-L icosphere (20 * 4^L triangles, L = 8 -> 1,310,720)
//              displaced by seeded 4-octave value noise (amplitude 0.08) on a
//              2-triangle ground; per-region (icosahedron face) gold GGX with
//              roughness U[0.05, 0.5]; quad area light + point light;
//              1024^2, 64 spp, depth 8
//   caustics   level-6 glass icosphere (eta 1.5) on a matte plane, 30 degree
//              spot light; SPPM 512^2, 64 x 2^20 photons, r0 0.01, depth 10
//   batch_job  one job of a batched stream: displaced icosphere of a requested
//              size with a material from {Lambert, gold GGX, dielectric}
//   materials  the cornell box with every material kind: Oren-Nayar walls,
//              plastic and mirror boxes, an uber sphere (all five lobes), a
//              rough glass sphere and a gold sphere; ceiling light + point
//              light; 128^2, 16 spp, depth 6
//   sky        spheres (mirror, glass, uber, plastic) on an Oren-Nayar ground
//              under a 64 x 32 sky envmap (sun disc, worldUp +y, turned 30
//              degrees about its up axis) and a distant sun; 128^2, 16 spp,
//              depth 6
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "host_math.h"
#include "host_scene.h"

namespace orion_amd {
// createTransform (ori_loader.cpp)
bool create_transform(const std::vector<std::pair<std::string, std::string>> &pv, Xform *out, std::string *err);
namespace {

struct Rng {  // splitmix64
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct V { double x, y, z; };
inline V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V vmul(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V vcross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V vnorm(V a) { double l = std::sqrt(vdot(a, a)); return vmul(a, 1.0 / l); }

struct Builder {
  HostScene *s;
  // appends a triangle; flips the winding so the geometric normal agrees with n
  void tri(V a, V b, V c, V na, V nb, V nc, int mat, int light) {
    V g = vcross(vsub(b, a), vsub(c, a));
    V navg = vadd(vadd(na, nb), nc);
    if (vdot(g, navg) < 0) { std::swap(b, c); std::swap(nb, nc); }
    const V pts[6] = {a, b, c, na, nb, nc};
    for (const V &p : pts) {
      s->tri_pn.push_back((float)p.x);
      s->tri_pn.push_back((float)p.y);
      s->tri_pn.push_back((float)p.z);
    }
    s->tri_mat.push_back(mat);
    s->tri_light.push_back(light);
  }
  void quad(V a, V b, V c, V d, V n, int mat, const float *Le) {
    for (int k = 0; k < 2; ++k) {
      int light = -1;
      if (Le) {
        light = s->n_lights();
        float rec[32];
        area_light_record(s->n_tris(), Le, rec);
        s->lights.insert(s->lights.end(), rec, rec + 32);
      }
      if (k == 0) tri(a, b, c, n, n, n, mat, light);
      else tri(a, c, d, n, n, n, mat, light);
    }
  }
  int material(const float *rec) {
    int m = s->n_mats();
    s->mats.insert(s->mats.end(), rec, rec + kMatFloats);
    return m;
  }
  void box(V center, V half, double angle_deg, int mat) {
    double a = angle_deg * M_PI / 180.0, ca = std::cos(a), sa = std::sin(a);
    auto xf = [&](V p) { return V{center.x + ca * p.x + sa * p.z, center.y + p.y, center.z - sa * p.x + ca * p.z}; };
    auto xn = [&](V n) { return V{ca * n.x + sa * n.z, n.y, -sa * n.x + ca * n.z}; };
    const double h[3] = {half.x, half.y, half.z};
    for (int a = 0; a < 3; ++a)
      for (int sg = -1; sg <= 1; sg += 2) {
        int bb = (a + 1) % 3, cc = (a + 2) % 3;
        auto corner = [&](int s1, int s2) {
          double q[3] = {0, 0, 0};
          q[a] = sg * h[a];
          q[bb] = s1 * h[bb];
          q[cc] = s2 * h[cc];
          return xf(V{q[0], q[1], q[2]});
        };
        double nn[3] = {0, 0, 0};
        nn[a] = sg;
        quad(corner(-1, -1), corner(1, -1), corner(1, 1), corner(-1, 1), xn(V{nn[0], nn[1], nn[2]}), mat,
             nullptr);
      }
  }
};

void set_camera(HostScene *s, const float la[9], float fov, int W, int H) {
  std::memcpy(s->lookat, la, sizeof(float) * 9);
  s->fov = fov;
  camera_record(la, fov, W, H, s->camera);
}

// ---------------------------------------------------------------- cornell
void gen_cornell(HostScene *s, int W, int H) {
  Builder b{s};
  float grey[kMatFloats], red[kMatFloats], green[kMatFloats], black[kMatFloats];
  matte_record(0.75f, 0.75f, 0.75f, grey);
  matte_record(0.63f, 0.06f, 0.05f, red);
  matte_record(0.14f, 0.45f, 0.09f, green);
  matte_record(0.f, 0.f, 0.f, black);
  int mg = b.material(grey), mr = b.material(red), mgr = b.material(green), mb = b.material(black);
  const float Le[3] = {15.f, 15.f, 15.f};
  // light first: its two triangles are lights 0 and 1
  b.quad({0.375, 0.999, 0.375}, {0.625, 0.999, 0.375}, {0.625, 0.999, 0.625}, {0.375, 0.999, 0.625},
         {0, -1, 0}, mb, Le);
  b.quad({0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, mg, nullptr);       // floor
  b.quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}, {0, -1, 0}, mg, nullptr);      // ceiling
  b.quad({0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, mg, nullptr);       // back
  b.quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}, mr, nullptr);       // left
  b.quad({1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {-1, 0, 0}, mgr, nullptr);     // right
  b.box({0.33, 0.15, 0.62}, {0.15, 0.15, 0.15}, -17.0, mg);                          // short box
  b.box({0.66, 0.30, 0.36}, {0.15, 0.30, 0.15}, 20.0, mg);                           // tall box
  const float la[9] = {0.5f, 0.5f, 2.4f, 0.5f, 0.5f, 0.0f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 0;
  s->info.spp = 16;
  s->info.max_depth = 5;
}

// ---------------------------------------------------------------- materials
void add_sphere(Builder &b, int L, V center, double radius, double amp, uint64_t noise_seed,
                const std::vector<int> &region_mat);

void gen_materials(HostScene *s, int W, int H) {
  Builder b{s};
  float rec[kMatFloats];
  matte_record(0.75f, 0.75f, 0.75f, rec, 20.f);  // Oren-Nayar
  int mwall = b.material(rec);
  matte_record(0.63f, 0.06f, 0.05f, rec, 35.f);
  int mred = b.material(rec);
  matte_record(0.14f, 0.45f, 0.09f, rec);         // Lambert
  int mgreen = b.material(rec);
  matte_record(0.f, 0.f, 0.f, rec);
  int mblack = b.material(rec);
  const float pkd[3] = {0.25f, 0.3f, 0.65f}, pks[3] = {0.4f, 0.4f, 0.4f};
  plastic_record(pkd, pks, 0.12f, rec);
  int mplastic = b.material(rec);
  const float mkr[3] = {0.9f, 0.85f, 0.8f};
  mirror_record(mkr, rec);
  int mmirror = b.material(rec);
  const float ukd[3] = {0.5f, -0.1f, 0.3f}, uks[3] = {0.2f, 0.2f, 0.2f}, ukr[3] = {0.25f, 0.25f, 0.25f},
              ukt[3] = {0.4f, 0.5f, 0.6f}, uop[3] = {0.6f, 0.75f, 1.2f};
  uber_record(ukd, uks, ukr, ukt, uop, 0.2f, 1.45f, rec);
  int muber = b.material(rec);
  const float gkr[3] = {0.95f, 0.95f, 0.95f}, gkt[3] = {0.9f, 0.9f, 0.9f};
  glass_record(gkr, gkt, 0.18f, 1.5f, rec);
  int mglass = b.material(rec);
  float eta[3], k[3];
  metal_preset("gold", eta, k);
  metal_record(eta, k, 0.25f, rec);
  int mgold = b.material(rec);
  const float Le[3] = {15.f, 15.f, 15.f};
  b.quad({0.375, 0.999, 0.375}, {0.625, 0.999, 0.375}, {0.625, 0.999, 0.625}, {0.375, 0.999, 0.625},
         {0, -1, 0}, mblack, Le);
  b.quad({0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, mwall, nullptr);     // floor
  b.quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}, {0, -1, 0}, mwall, nullptr);    // ceiling
  b.quad({0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, mwall, nullptr);     // back
  b.quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}, mred, nullptr);      // left
  b.quad({1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {-1, 0, 0}, mgreen, nullptr);   // right
  b.box({0.30, 0.12, 0.70}, {0.12, 0.12, 0.12}, -17.0, mplastic);                     // short box
  b.box({0.70, 0.28, 0.30}, {0.12, 0.28, 0.12}, 20.0, mmirror);                       // tall box
  add_sphere(b, 3, {0.30, 0.38, 0.68}, 0.13, 0.0, 0, std::vector<int>(1, muber));
  add_sphere(b, 3, {0.72, 0.12, 0.72}, 0.12, 0.0, 0, std::vector<int>(1, mglass));
  add_sphere(b, 2, {0.18, 0.10, 0.25}, 0.10, 0.0, 0, std::vector<int>(1, mgold));
  const float I[3] = {0.6f, 0.6f, 0.5f}, from[3] = {0.5f, 0.8f, 0.9f};
  float lrec[32];
  point_light_record(I, from, lrec);
  s->lights.insert(s->lights.end(), lrec, lrec + 32);
  const float la[9] = {0.5f, 0.5f, 2.4f, 0.5f, 0.5f, 0.0f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 0;
  s->info.spp = 16;
  s->info.max_depth = 6;
}

// ---------------------------------------------------------------- sky
void gen_sky(HostScene *s, int W, int H) {
  Builder b{s};
  float rec[kMatFloats];
  matte_record(0.5f, 0.45f, 0.4f, rec, 10.f);
  int mground = b.material(rec);
  const float kr[3] = {0.95f, 0.95f, 0.95f};
  mirror_record(kr, rec);
  int mmirror = b.material(rec);
  const float gk[3] = {1.f, 1.f, 1.f};
  glass_record(gk, gk, 0.f, 1.5f, rec);
  int mglass = b.material(rec);
  const float ukd[3] = {0.6f, 0.3f, 0.2f}, uks[3] = {0.3f, 0.3f, 0.3f}, ukr[3] = {0.2f, 0.2f, 0.2f},
              ukt[3] = {0.3f, 0.3f, 0.3f}, uop[3] = {0.8f, 0.8f, 0.8f};
  uber_record(ukd, uks, ukr, ukt, uop, 0.15f, 1.4f, rec);
  int muber = b.material(rec);
  const float pkd[3] = {0.1f, 0.5f, 0.2f}, pks[3] = {0.5f, 0.5f, 0.5f};
  plastic_record(pkd, pks, 0.05f, rec);
  int mplastic = b.material(rec);
  b.quad({-4, 0, -4}, {4, 0, -4}, {4, 0, 4}, {-4, 0, 4}, {0, 1, 0}, mground, nullptr);
  add_sphere(b, 3, {-1.1, 0.5, 0.0}, 0.5, 0.0, 0, std::vector<int>(1, mmirror));
  add_sphere(b, 3, {0.0, 0.5, 0.6}, 0.5, 0.0, 0, std::vector<int>(1, mglass));
  add_sphere(b, 3, {1.1, 0.5, 0.0}, 0.5, 0.0, 0, std::vector<int>(1, muber));
  add_sphere(b, 2, {0.4, 0.25, 1.45}, 0.25, 0.0, 0, std::vector<int>(1, mplastic));
  // distant sun
  const float sunL[3] = {2.0f, 1.9f, 1.7f}, from[3] = {0.f, 0.f, 0.f}, to[3] = {0.4f, 1.f, 0.3f};
  float lrec[32];
  distant_light_record(sunL, from, to, lrec);
  s->lights.insert(s->lights.end(), lrec, lrec + 32);
  // envmap: 8-bit texels (byte / 255), theta from the map's +z (= worldUp)
  EnvImage img;
  img.width = 64;
  img.height = 32;
  img.file = "";
  img.xform = {{"rotate", "0,0,1,30"}};
  img.rgb.resize(64 * 32 * 3);
  const float inv255 = 0.0039215686274509803921568627451f;
  for (int j = 0; j < 32; ++j)
    for (int i = 0; i < 64; ++i) {
      double u = (i + 0.5) / 64, v = (j + 0.5) / 32;
      double c[3];
      if (v < 0.5) {
        c[0] = 0.3 + 0.4 * (1 - 2 * v); c[1] = 0.5 + 0.3 * (1 - 2 * v); c[2] = 0.9;
      } else {
        c[0] = 0.25; c[1] = 0.2; c[2] = 0.15;
      }
      double du = u - 0.3, dv = v - 0.22;
      if (du * du + dv * dv < 0.05 * 0.05) c[0] = c[1] = c[2] = 1.0;
      for (int k = 0; k < 3; ++k) img.rgb[3 * (j * 64 + i) + k] = (float)(int)std::lround(c[k] * 255) * inv255;
    }
  Xform t;
  std::string err;
  create_transform(img.xform, &t, &err);
  const float up[3] = {0.f, 1.f, 0.f}, envL[3] = {1.5f, 1.5f, 1.5f};
  float l2w[16], w2l[16];
  envmap_orientation(&t.m.m[0][0], &t.mInv.m[0][0], up, l2w, w2l);
  envmap_light_record((int)s->env.size(), envL, 64, 32, 1, l2w, w2l, up, lrec);
  s->lights.insert(s->lights.end(), lrec, lrec + 32);
  s->env.push_back(std::move(img));
  const float la[9] = {0.f, 1.2f, 4.0f, 0.f, 0.5f, 0.f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 0;
  s->info.spp = 16;
  s->info.max_depth = 6;
}

// ---------------------------------------------------------------- value noise
struct Noise {
  uint64_t seed;
  double lattice(int64_t x, int64_t y, int64_t z) const {
    uint64_t h = seed ^ ((uint64_t)x * 0x9E3779B185EBCA87ull) ^ ((uint64_t)y * 0xC2B2AE3D27D4EB4Full) ^
                 ((uint64_t)z * 0x165667B19E3779F9ull);
    h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
    return (h >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // [-1, 1)
  }
  double value(V p) const {
    double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    auto sm = [](double t) { return t * t * (3 - 2 * t); };
    tx = sm(tx); ty = sm(ty); tz = sm(tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d);
    auto l = [](double a, double b, double t) { return a + (b - a) * t; };
    double x00 = l(c[0][0][0], c[1][0][0], tx), x10 = l(c[0][1][0], c[1][1][0], tx);
    double x01 = l(c[0][0][1], c[1][0][1], tx), x11 = l(c[0][1][1], c[1][1][1], tx);
    return l(l(x00, x10, ty), l(x01, x11, ty), tz);
  }
  double fbm(V p, int octaves) const {  // lacunarity 2, gain 1/2, normalised to [-1, 1]
    double sum = 0, amp = 1, norm = 0, f = 1;
    for (int o = 0; o < octaves; ++o) {
      sum += amp * value(vmul(p, f));
      norm += amp;
      amp *= 0.5;
      f *= 2;
    }
    return sum / norm;
  }
};

// icosphere: 12-vertex icosahedron, L midpoint subdivisions, unit radius.
// Outputs vertices and faces (with the base-face region of each face).
void icosphere(int L, std::vector<V> *verts, std::vector<int> *faces, std::vector<int> *region) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<V> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto &p : v) p = vnorm(p);
  std::vector<int> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                        11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                        3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
  std::vector<int> reg(20);
  for (int i = 0; i < 20; ++i) reg[i] = i;
  for (int lv = 0; lv < L; ++lv) {
    std::map<uint64_t, int> mid;
    auto midpoint = [&](int a, int b) {
      uint64_t key = a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a;
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      v.push_back(vnorm(vmul(vadd(v[a], v[b]), 0.5)));
      int id = (int)v.size() - 1;
      mid.emplace(key, id);
      return id;
    };
    std::vector<int> nf;
    std::vector<int> nr;
    nf.reserve(f.size() * 4);
    nr.reserve(reg.size() * 4);
    for (size_t i = 0; i < f.size(); i += 3) {
      int a = f[i], b = f[i + 1], c = f[i + 2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      int r = reg[i / 3];
      const int tris[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
      nf.insert(nf.end(), tris, tris + 12);
      nr.insert(nr.end(), {r, r, r, r});
    }
    f.swap(nf);
    reg.swap(nr);
  }
  *verts = std::move(v);
  *faces = std::move(f);
  *region = std::move(reg);
}

// Appends a (possibly displaced) icosphere with area-weighted vertex normals.
void add_sphere(Builder &b, int L, V center, double radius, double amp, uint64_t noise_seed,
                const std::vector<int> &region_mat) {
  std::vector<V> v;
  std::vector<int> f, reg;
  icosphere(L, &v, &f, &reg);
  Noise nz{noise_seed};
  std::vector<V> p(v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    double r = radius * (1.0 + (amp > 0 ? amp * nz.fbm(vmul(v[i], 2.0), 4) : 0.0));
    p[i] = vadd(center, vmul(v[i], r));
  }
  std::vector<V> n(v.size(), V{0, 0, 0});
  if (amp > 0) {
    for (size_t i = 0; i < f.size(); i += 3) {
      V g = vcross(vsub(p[f[i + 1]], p[f[i]]), vsub(p[f[i + 2]], p[f[i]]));  // |g| = 2 * area
      for (int k = 0; k < 3; ++k) n[f[i + k]] = vadd(n[f[i + k]], g);
    }
    for (auto &q : n) q = vnorm(q);
  } else {
    n = v;  // smooth sphere: exact radial normals
  }
  b.s->tri_pn.reserve(b.s->tri_pn.size() + f.size() * 6);
  for (size_t i = 0; i < f.size(); i += 3) {
    int m = region_mat[reg[i / 3] % region_mat.size()];
    b.tri(p[f[i]], p[f[i + 1]], p[f[i + 2]], n[f[i]], n[f[i + 1]], n[f[i + 2]], m, -1);
  }
}

void gen_icosphere(HostScene *s, uint64_t seed, int L, int W, int H) {
  Builder b{s};
  Rng rng(seed);
  float eta[3], k[3];
  metal_preset("gold", eta, k);
  std::vector<int> region(20);
  for (int r = 0; r < 20; ++r) {
    float rec[kMatFloats];
    metal_record(eta, k, (float)(0.05 + 0.45 * rng.uniform()), rec);
    region[r] = b.material(rec);
  }
  float ground[kMatFloats], black[kMatFloats];
  matte_record(0.5f, 0.5f, 0.5f, ground);
  matte_record(0.f, 0.f, 0.f, black);
  int mground = b.material(ground), mblack = b.material(black);
  const float Le[3] = {8.f, 8.f, 8.f};
  b.quad({-0.5, 4.0, -0.5}, {0.5, 4.0, -0.5}, {0.5, 4.0, 0.5}, {-0.5, 4.0, 0.5}, {0, -1, 0}, mblack, Le);
  add_sphere(b, L, {0.0, 1.08, 0.0}, 1.0, 0.08, seed ^ 0x5eedull, region);
  b.quad({-10, 0, -10}, {10, 0, -10}, {10, 0, 10}, {-10, 0, 10}, {0, 1, 0}, mground, nullptr);
  float rec[32];
  const float I[3] = {12.f, 12.f, 12.f}, from[3] = {3.f, 3.5f, 3.f};
  point_light_record(I, from, rec);
  s->lights.insert(s->lights.end(), rec, rec + 32);
  const float la[9] = {0.f, 1.6f, 4.2f, 0.f, 0.9f, 0.f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 0;
  s->info.spp = 64;
  s->info.max_depth = 8;
}

void gen_caustics(HostScene *s, int L, int W, int H) {
  Builder b{s};
  float glass[kMatFloats], plane[kMatFloats];
  const float one[3] = {1.f, 1.f, 1.f};
  glass_record(one, one, 0.f, 1.5f, glass);
  matte_record(0.5f, 0.5f, 0.5f, plane);
  int mglass = b.material(glass), mplane = b.material(plane);
  add_sphere(b, L, {0.0, 0.5, 0.0}, 0.5, 0.0, 0, std::vector<int>(1, mglass));
  b.quad({-5, 0, -5}, {5, 0, -5}, {5, 0, 5}, {-5, 0, 5}, {0, 1, 0}, mplane, nullptr);
  float rec[32];
  const float I[3] = {60.f, 60.f, 60.f}, from[3] = {1.2f, 2.6f, 0.4f}, to[3] = {0.f, 0.5f, 0.f};
  spot_light_record(I, from, to, 30.f, 0.f, rec);
  s->lights.insert(s->lights.end(), rec, rec + 32);
  const float la[9] = {0.f, 2.0f, 2.8f, -0.2f, 0.25f, -0.3f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 1;
  s->info.max_depth = 10;
  s->info.sppm_iterations = 64;
  s->info.sppm_photons = 1 << 20;
  s->info.sppm_radius = 0.01f;
}

// One job of the batched / served streams: icosphere level from the triangle
// budget, material mix 0 Lambert, 1 gold GGX, 2 dielectric, 3 = drawn.
void gen_batch_job(HostScene *s, uint64_t seed, double tris, int W, int H, int mix) {
  Builder b{s};
  Rng rng(seed);
  int L = 1;
  while (20.0 * std::pow(4.0, L + 1) <= tris && L < 9) ++L;
  if (mix == 3) mix = (int)(rng.uniform() * 3) % 3;
  float rec[kMatFloats];
  if (mix == 0) {
    matte_record((float)(0.3 + 0.6 * rng.uniform()), (float)(0.3 + 0.6 * rng.uniform()),
                 (float)(0.3 + 0.6 * rng.uniform()), rec);
  } else if (mix == 1) {
    float eta[3], k[3];
    metal_preset("gold", eta, k);
    metal_record(eta, k, (float)(0.05 + 0.45 * rng.uniform()), rec);
  } else {
    const float one[3] = {1.f, 1.f, 1.f};
    glass_record(one, one, 0.f, 1.5f, rec);
  }
  int mobj = b.material(rec);
  float ground[kMatFloats], black[kMatFloats];
  matte_record(0.5f, 0.5f, 0.5f, ground);
  matte_record(0.f, 0.f, 0.f, black);
  int mground = b.material(ground), mblack = b.material(black);
  const float Le[3] = {8.f, 8.f, 8.f};
  b.quad({-0.5, 4.0, -0.5}, {0.5, 4.0, -0.5}, {0.5, 4.0, 0.5}, {-0.5, 4.0, 0.5}, {0, -1, 0}, mblack, Le);
  add_sphere(b, L, {0.0, 1.08, 0.0}, 1.0, 0.08, seed ^ 0xabcdull, std::vector<int>(1, mobj));
  b.quad({-10, 0, -10}, {10, 0, -10}, {10, 0, 10}, {-10, 0, 10}, {0, 1, 0}, mground, nullptr);
  float lrec[32];
  const float I[3] = {12.f, 12.f, 12.f}, from[3] = {3.f, 3.5f, 3.f};
  point_light_record(I, from, lrec);
  s->lights.insert(s->lights.end(), lrec, lrec + 32);
  const float la[9] = {0.f, 1.6f, 4.2f, 0.f, 0.9f, 0.f, 0.f, 1.f, 0.f};
  set_camera(s, la, 40.f, W, H);
  s->info.integrator = 0;
  s->info.spp = 16;
  s->info.max_depth = 8;
}

}  // namespace

int generate_scene(const std::string &name, uint64_t seed, const double *params, int n,
                   HostScene *out, std::string *err) {
  auto P = [&](int i, double d) { return (params && i < n) ? params[i] : d; };
  HostScene s;
  if (name == "cornell") {
    gen_cornell(&s, (int)P(0, 256), (int)P(1, 256));
  } else if (name == "icosphere") {
    int L = (int)P(0, 8);
    if (L < 0 || L > 9) { *err = "icosphere level must be in [0, 9]"; return ORION_ERR_ARG; }
    gen_icosphere(&s, seed, L, (int)P(1, 1024), (int)P(2, 1024));
  } else if (name == "caustics") {
    int L = (int)P(0, 6);
    if (L < 0 || L > 9) { *err = "caustics level must be in [0, 9]"; return ORION_ERR_ARG; }
    gen_caustics(&s, L, (int)P(1, 512), (int)P(2, 512));
  } else if (name == "sky") {
    gen_sky(&s, (int)P(0, 128), (int)P(1, 128));
  } else if (name == "materials") {
    gen_materials(&s, (int)P(0, 128), (int)P(1, 128));
  } else if (name == "batch_job") {
    gen_batch_job(&s, seed, P(0, 20000), (int)P(1, 256), (int)P(2, 256), (int)P(3, 3));
  } else {
    *err = "unknown scene '" + name + "'";
    return ORION_ERR_ARG;
  }
  s.refresh_info();
  *out = std::move(s);
  return ORION_OK;
}

}  // namespace orion_amd
