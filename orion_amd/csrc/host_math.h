// Host-side transform math used when resolving a scene (camera, lights,
// model transforms).  The operation order matches core/transform.cpp so the
// resolved records equal what orion computes for the same .ori input.
#pragma once
#include <cmath>
#include <cstring>
#include <utility>

namespace orion_amd {

struct Mat4 {
  float m[4][4];
  static Mat4 identity() {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
    return r;
  }
  static Mat4 make(float t00, float t01, float t02, float t03, float t10, float t11, float t12,
                   float t13, float t20, float t21, float t22, float t23, float t30, float t31,
                   float t32, float t33) {
    Mat4 r;
    float v[16] = {t00, t01, t02, t03, t10, t11, t12, t13, t20, t21, t22, t23, t30, t31, t32, t33};
    std::memcpy(r.m, v, sizeof(v));
    return r;
  }
};

// transform.cpp:29-35
inline Mat4 transpose(const Mat4 &a) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

// transform.cpp:36-44 (sum evaluated left to right)
inline Mat4 mul(const Mat4 &a, const Mat4 &b) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                  a.m[i][3] * b.m[3][j];
  return r;
}

// Gauss-Jordan with full pivoting, transform.cpp:45-99 (same pivot rule,
// same operation order, so results match bit for bit).
inline Mat4 inverse(const Mat4 &a) {
  int indxc[4], indxr[4];
  int ipiv[4] = {0, 0, 0, 0};
  float minv[4][4];
  std::memcpy(minv, a.m, sizeof(minv));
  for (int i = 0; i < 4; i++) {
    int irow = 0, icol = 0;
    float big = 0.f;
    for (int j = 0; j < 4; j++) {
      if (ipiv[j] != 1) {
        for (int k = 0; k < 4; k++) {
          if (ipiv[k] == 0) {
            if (std::abs(minv[j][k]) >= big) {
              big = float(std::abs(minv[j][k]));
              irow = j;
              icol = k;
            }
          }
        }
      }
    }
    ++ipiv[icol];
    if (irow != icol)
      for (int k = 0; k < 4; ++k) std::swap(minv[irow][k], minv[icol][k]);
    indxr[i] = irow;
    indxc[i] = icol;
    float pivinv = 1.f / minv[icol][icol];
    minv[icol][icol] = 1.;
    for (int j = 0; j < 4; j++) minv[icol][j] *= pivinv;
    for (int j = 0; j < 4; j++) {
      if (j != icol) {
        float save = minv[j][icol];
        minv[j][icol] = 0;
        for (int k = 0; k < 4; k++) minv[j][k] -= minv[icol][k] * save;
      }
    }
  }
  for (int j = 3; j >= 0; j--) {
    if (indxr[j] != indxc[j]) {
      for (int k = 0; k < 4; k++) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
    }
  }
  Mat4 r;
  std::memcpy(r.m, minv, sizeof(minv));
  return r;
}

struct Xform {
  Mat4 m, mInv;
  Xform() : m(Mat4::identity()), mInv(Mat4::identity()) {}
  Xform(const Mat4 &a) : m(a), mInv(inverse(a)) {}
  Xform(const Mat4 &a, const Mat4 &b) : m(a), mInv(b) {}
};
inline Xform inverse(const Xform &t) { return Xform(t.mInv, t.m); }
inline Xform operator*(const Xform &a, const Xform &b) {  // transform.cpp:110-113
  return Xform(mul(a.m, b.m), mul(b.mInv, a.mInv));
}

struct F3 { float x, y, z; };
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) {
  return f3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
inline F3 normalize(F3 a) {
  float inv = 1.0f / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return f3(a.x * inv, a.y * inv, a.z * inv);
}
// geometry.h:866-868
inline float radians(float deg) { return deg * static_cast<float>(0.01745329251994329576923690768489); }

// transform.h:118-131
inline F3 apply_point(const Mat4 &m, F3 p) {
  float x = p.x, y = p.y, z = p.z;
  float xp = m.m[0][0] * x + m.m[0][1] * y + m.m[0][2] * z + m.m[0][3];
  float yp = m.m[1][0] * x + m.m[1][1] * y + m.m[1][2] * z + m.m[1][3];
  float zp = m.m[2][0] * x + m.m[2][1] * y + m.m[2][2] * z + m.m[2][3];
  float wp = m.m[3][0] * x + m.m[3][1] * y + m.m[3][2] * z + m.m[3][3];
  if (wp == 1) return f3(xp, yp, zp);
  float inv = (float)1 / wp;
  return f3(inv * xp, inv * yp, inv * zp);
}
inline F3 apply_vector(const Mat4 &m, F3 v) {  // transform.h:133-139
  return f3(m.m[0][0] * v.x + m.m[0][1] * v.y + m.m[0][2] * v.z,
            m.m[1][0] * v.x + m.m[1][1] * v.y + m.m[1][2] * v.z,
            m.m[2][0] * v.x + m.m[2][1] * v.y + m.m[2][2] * v.z);
}
inline F3 apply_normal(const Xform &t, F3 n) {  // transform.h:160-167 (uses mInv^T)
  const Mat4 &mi = t.mInv;
  return f3(mi.m[0][0] * n.x + mi.m[1][0] * n.y + mi.m[2][0] * n.z,
            mi.m[0][1] * n.x + mi.m[1][1] * n.y + mi.m[2][1] * n.z,
            mi.m[0][2] * n.x + mi.m[1][2] * n.y + mi.m[2][2] * n.z);
}

inline Xform translate(F3 d) {  // transform.cpp:132-139
  Mat4 m = Mat4::make(1, 0, 0, d.x, 0, 1, 0, d.y, 0, 0, 1, d.z, 0, 0, 0, 1);
  Mat4 mi = Mat4::make(1, 0, 0, -d.x, 0, 1, 0, -d.y, 0, 0, 1, -d.z, 0, 0, 0, 1);
  return Xform(m, mi);
}
inline Xform scale(float x, float y, float z) {  // transform.cpp:140-145
  Mat4 m = Mat4::make(x, 0, 0, 0, 0, y, 0, 0, 0, 0, z, 0, 0, 0, 0, 1);
  Mat4 mi = Mat4::make(1 / x, 0, 0, 0, 0, 1 / y, 0, 0, 0, 0, 1 / z, 0, 0, 0, 0, 1);
  return Xform(m, mi);
}
inline Xform rotate(float theta, F3 axis) {  // transform.cpp:174-196
  F3 a = normalize(axis);
  float sinTheta = std::sin(radians(theta));
  float cosTheta = std::cos(radians(theta));
  Mat4 m = Mat4::identity();
  m.m[0][0] = a.x * a.x + (1 - a.x * a.x) * cosTheta;
  m.m[0][1] = a.x * a.y * (1 - cosTheta) - a.z * sinTheta;
  m.m[0][2] = a.x * a.z * (1 - cosTheta) + a.y * sinTheta;
  m.m[0][3] = 0;
  m.m[1][0] = a.x * a.y * (1 - cosTheta) + a.z * sinTheta;
  m.m[1][1] = a.y * a.y + (1 - a.y * a.y) * cosTheta;
  m.m[1][2] = a.y * a.z * (1 - cosTheta) - a.x * sinTheta;
  m.m[1][3] = 0;
  m.m[2][0] = a.x * a.z * (1 - cosTheta) - a.y * sinTheta;
  m.m[2][1] = a.y * a.z * (1 - cosTheta) + a.x * sinTheta;
  m.m[2][2] = a.z * a.z + (1 - a.z * a.z) * cosTheta;
  m.m[2][3] = 0;
  return Xform(m, transpose(m));
}
inline Xform look_at(F3 pos, F3 look, F3 up) {  // transform.cpp:197-217
  F3 f = normalize(sub(look, pos));
  F3 s = normalize(cross(f, up));
  F3 u = cross(s, f);
  Mat4 r = Mat4::identity();
  r.m[0][0] = s.x; r.m[0][1] = s.y; r.m[0][2] = s.z;
  r.m[1][0] = u.x; r.m[1][1] = u.y; r.m[1][2] = u.z;
  r.m[2][0] = -f.x; r.m[2][1] = -f.y; r.m[2][2] = -f.z;
  r.m[0][3] = -dot(s, pos);
  r.m[1][3] = -dot(u, pos);
  r.m[2][3] = dot(f, pos);
  return Xform(inverse(r), r);
}
// geometry.h:902-909
inline void coordinate_system(F3 z, F3 *x, F3 *y) {
  if (std::abs(z.x) > std::abs(z.y)) {
    float inv = 1.0f / std::sqrt(z.x * z.x + z.z * z.z);
    *y = f3(-z.z * inv, 0 * inv, z.x * inv);
  } else {
    float inv = 1.0f / std::sqrt(z.y * z.y + z.z * z.z);
    *y = f3(0 * inv, z.z * inv, -z.y * inv);
  }
  *x = cross(*y, z);
}

}  // namespace orion_amd
