// Device-side shading arithmetic for the gfx950 path (float, no FMA
// contraction: the library is compiled with -ffp-contract=off).  Every
// function follows the reference operation by operation -- association
// order, v / s as v * (1 / s) for vectors but true division for spectra,
// std::min/std::max argument order -- so a GPU sample path is the same
// float sequence as orion's.  References are to the reference's src/.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orion_amd {
namespace dev {

constexpr float kEps = 0.000001f;            // orion.h:48
constexpr float kShadowEps = 0.01f;          // orion.h:49
constexpr float kFltMax = 3.40282347e+38f;   // orion.h:51 fInfinity
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float kInv4Pi = 0.07957747154594766788f;
constexpr float kOneMinusEps = (float)0.99999994;  // bsdf.cpp:15
enum { REFL = 1, TRANS = 2, DIFF = 4, GLOSSY = 8, SPEC = 16, ALL = 31 };

__device__ __forceinline__ float mn(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ float mx(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ bool is_inf(float x) { return __builtin_isinf(x); }

// sin / cos of a float argument evaluated in FP64 and rounded once: the
// correctly rounded float except in ~1e-8 of inputs (the CPU side does the
// same with libm's double routines, so sample paths agree).
__device__ __forceinline__ float dsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float dcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float dacos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float datan2(float y, float x) { return (float)atan2((double)y, (double)x); }

// 1 / x rounded to nearest, as IEEE division rounds it: the hardware
// reciprocal estimate refined by one Newton-Raphson step with fused
// multiply-adds (Markstein's correction r + r (1 - x r)).  For |x| in
// [2^-125, 2^125) the result equals the IEEE quotient bit for bit on gfx950
// -- every float in that range is compared by orion_selfcheck_rcp -- and
// other magnitudes (zero, infinities, NaN, the extremes) take the division.
__device__ __forceinline__ float rcp_rn(float x) {
  const float ax = fabsf(x);
  if (__builtin_expect(ax >= 0x1p-125f && ax < 0x1p125f, 1)) {
    const float r = __builtin_amdgcn_rcpf(x);
    return fmaf(fmaf(-x, r, 1.0f), r, r);
  }
  return 1.0f / x;
}

struct V3 { float x, y, z; };
__device__ __forceinline__ V3 mk(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 divs(V3 a, float f) { float inv = 1.0f / f; return mk(a.x * inv, a.y * inv, a.z * inv); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float absdot(V3 a, V3 b) { return fabsf(a.x * b.x + a.y * b.y + a.z * b.z); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__device__ __forceinline__ float lensq(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ V3 normalize(V3 a) { return divs(a, sqrtf(lensq(a))); }

struct Sp { float r, g, b; };
__device__ __forceinline__ Sp sp(float v) { Sp s; s.r = v; s.g = v; s.b = v; return s; }
__device__ __forceinline__ Sp sp3(float r, float g, float b) { Sp s; s.r = r; s.g = g; s.b = b; return s; }
__device__ __forceinline__ Sp operator+(Sp a, Sp b) { return sp3(b.r + a.r, b.g + a.g, b.b + a.b); }
__device__ __forceinline__ Sp operator-(Sp a, Sp b) { return sp3(a.r - b.r, a.g - b.g, a.b - b.b); }
__device__ __forceinline__ Sp operator*(Sp a, Sp b) { return sp3(b.r * a.r, b.g * a.g, b.b * a.b); }
__device__ __forceinline__ Sp operator/(Sp a, Sp b) { return sp3(a.r / b.r, a.g / b.g, a.b / b.b); }
__device__ __forceinline__ Sp operator*(Sp a, float f) { return sp3(a.r * f, a.g * f, a.b * f); }
__device__ __forceinline__ Sp operator/(Sp a, float f) { return sp3(a.r / f, a.g / f, a.b / f); }
__device__ __forceinline__ Sp operator+(Sp a, float f) { return sp3(a.r + f, a.g + f, a.b + f); }
__device__ __forceinline__ Sp operator-(Sp a, float f) { return sp3(a.r - f, a.g - f, a.b - f); }
__device__ __forceinline__ void acc(Sp &a, Sp b) { a.r += b.r; a.g += b.g; a.b += b.b; }
__device__ __forceinline__ void mulby(Sp &a, Sp b) { a.r *= b.r; a.g *= b.g; a.b *= b.b; }
__device__ __forceinline__ bool black(Sp a) { return !(a.r != 0.0f || a.g != 0.0f || a.b != 0.0f); }
__device__ __forceinline__ float maxc(Sp a) { return mx(a.r, mx(a.g, a.b)); }
__device__ __forceinline__ float intensity(Sp a) { return 0.212671f * a.r + 0.715160f * a.g + 0.072169f * a.b; }
__device__ __forceinline__ Sp ssqrt(Sp a) { return sp3(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }
__device__ __forceinline__ float clampf(float f, float lo, float hi) { return f < lo ? lo : (f > hi ? hi : f); }

// ---------------------------------------------------------------- sampling (sampling.cpp)
__device__ __forceinline__ V3 cosine_hemisphere(float u0, float u1) {
  float r = sqrtf(u0);
  float theta = 2 * kPi * u1;
  float x = r * dcos(theta), y = r * dsin(theta);
  float z = sqrtf(mx(0.0f, 1 - x * x - y * y));
  return mk(x, y, z);
}
__device__ __forceinline__ V3 uniform_sphere(float u0, float u1) {
  float z = 1 - 2 * u0;
  float r = sqrtf(mx(0.0f, 1.0f - z * z));
  float phi = 2 * kPi * u1;
  return mk(r * dcos(phi), r * dsin(phi), z);
}
__device__ __forceinline__ V3 uniform_cone(float u0, float u1, float cosThetaMax) {
  float cosTheta = (1.0f - u0) + u0 * cosThetaMax;
  float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
  float phi = u1 * 2 * kPi;
  return mk(dcos(phi) * sinTheta, dsin(phi) * sinTheta, cosTheta);
}
// sampling.cpp:35-54 (Shirley-Chiu concentric mapping)
__device__ __forceinline__ void concentric_disk(float u0, float u1, float *dx, float *dy) {
  float ox = 2.f * u0 - 1.0f, oy = 2.f * u1 - 1.0f;
  if (ox == 0 && oy == 0) { *dx = 0; *dy = 0; return; }
  float theta, r;
  if (fabsf(ox) > fabsf(oy)) {
    r = ox;
    theta = 0.78539816339744830961f * (oy / ox);
  } else {
    r = oy;
    theta = 1.57079632679489661923f - 0.78539816339744830961f * (ox / oy);
  }
  *dx = dcos(theta) * r;
  *dy = dsin(theta) * r;
}
// sampling.h:60-68
__device__ __forceinline__ float spherical_theta(V3 v) { return dacos(clampf(v.z, -1.0f, 1.0f)); }
__device__ __forceinline__ float spherical_phi(V3 v) {
  float p = datan2(v.y, v.x);
  return (p < 0) ? (p + 2 * kPi) : p;
}
__device__ __forceinline__ float power_heuristic(float fPdf, float gPdf) {
  float f = 1 * fPdf, g = 1 * gPdf;
  return (f * f) / (f * f + g * g);
}

// ---------------------------------------------------------------- ray spawning (geometry.h:824-863)
__device__ __forceinline__ float next_up(float v) {
  if (is_inf(v) && v > 0.f) return v;
  if (v == -0.f) v = 0.f;
  uint32_t ui = __float_as_uint(v);
  if (v >= 0) ++ui; else --ui;
  return __uint_as_float(ui);
}
__device__ __forceinline__ float next_down(float v) {
  if (is_inf(v) && v < 0.f) return v;
  if (v == 0.f) v = -0.f;
  uint32_t ui = __float_as_uint(v);
  if (v > 0) --ui; else ++ui;
  return __uint_as_float(ui);
}
__device__ __forceinline__ V3 offset_ray_origin(V3 p, V3 n, V3 w) {
  float maxValue = mx(mx(fabsf(p.x), fabsf(p.y)), fabsf(p.z));
  float c = maxValue * kEps;
  float d = fabsf(n.x) * c + fabsf(n.y) * c + fabsf(n.z) * c;
  V3 off = mul(n, d);
  if (dot(w, n) < 0) off = neg(off);
  V3 po = add(p, off);
  if (off.x > 0) po.x = next_up(po.x); else if (off.x < 0) po.x = next_down(po.x);
  if (off.y > 0) po.y = next_up(po.y); else if (off.y < 0) po.y = next_down(po.y);
  if (off.z > 0) po.z = next_up(po.z); else if (off.z < 0) po.z = next_down(po.z);
  return po;
}
__device__ __forceinline__ void coordinate_system(V3 z, V3 *x, V3 *y) {
  if (fabsf(z.x) > fabsf(z.y)) *y = divs(mk(-z.z, 0, z.x), sqrtf(z.x * z.x + z.z * z.z));
  else *y = divs(mk(0, z.z, -z.y), sqrtf(z.y * z.y + z.z * z.z));
  *x = cross(*y, z);
}
__device__ __forceinline__ V3 mat3(const float *m, V3 v) {
  return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// ---------------------------------------------------------------- BSDFs
enum { BX_LAMBERT = 0, BX_MF_REFL = 1, BX_MF_TRANS = 2, BX_FRESNEL_SPEC = 3, BX_OREN_NAYAR = 4,
       BX_SPEC_REFL = 5, BX_SPEC_TRANS = 6 };
enum { FR_CONDUCTOR = 0, FR_DIELECTRIC = 1, FR_NOOP = 2 };
// Load through a global-address-space pointer: scene arrays reach the
// kernels inside a job table in memory, where the compiler cannot prove
// they are global and would emit flat loads (which also hold the LDS
// counter, serialising the LDS stack behind outstanding fetches).
template <class T>
__device__ __forceinline__ T gld(const T *p) {
  typedef const __attribute__((address_space(1))) T gT;
  return *(gT *)p;
}
// vector types go through clang's native vectors (HIP's vector classes would
// copy through a generic reference)
typedef float gf4 __attribute__((ext_vector_type(4)));
typedef unsigned int gu4 __attribute__((ext_vector_type(4)));
typedef unsigned int gu2 __attribute__((ext_vector_type(2)));
template <>
__device__ __forceinline__ float4 gld<float4>(const float4 *p) {
  gf4 v = *(const __attribute__((address_space(1))) gf4 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
template <>
__device__ __forceinline__ int4 gld<int4>(const int4 *p) {
  gu4 v = *(const __attribute__((address_space(1))) gu4 *)p;
  return make_int4((int)v.x, (int)v.y, (int)v.z, (int)v.w);
}
template <>
__device__ __forceinline__ uint4 gld<uint4>(const uint4 *p) {
  gu4 v = *(const __attribute__((address_space(1))) gu4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
template <>
__device__ __forceinline__ uint2 gld<uint2>(const uint2 *p) {
  gu2 v = *(const __attribute__((address_space(1))) gu2 *)p;
  return make_uint2(v.x, v.y);
}
// three consecutive floats (one 12-byte load: the fourth word of a 16-byte
// record is not read)
__device__ __forceinline__ V3 gld3(const float *p) {
  typedef const __attribute__((address_space(1))) float gF;
  const gF *q = (const gF *)p;
  return mk(q[0], q[1], q[2]);
}


// One BxDF with its parameters (bsdf/*.h); A, B are Oren-Nayar's.
struct Lobe {
  int kind, type, fresnel, photon;
  Sp R, T, etaT, k;
  float alpha, etaA, etaB, A, B;
};

__device__ inline Sp fr_conductor(float cosThetaI, Sp etai, Sp etat, Sp k) {  // bsdf.cpp:124-147
  cosThetaI = clampf(cosThetaI, -1.0f, 1.0f);
  Sp eta = etat / etai;
  Sp etak = k / etai;
  float cos2 = cosThetaI * cosThetaI;
  float sin2 = 1.0f - cos2;
  Sp eta2 = eta * eta;
  Sp etak2 = etak * etak;
  Sp t0 = eta2 - etak2 - sin2;
  Sp a2plusb2 = ssqrt(t0 * t0 + eta2 * 4.0f * etak2);
  Sp t1 = a2plusb2 + cos2;
  Sp a = ssqrt((a2plusb2 + t0) * 0.5f);
  Sp t2 = a * (2.0f * cosThetaI);
  Sp Rs = (t1 - t2) / (t1 + t2);
  Sp t3 = a2plusb2 * cos2 + sin2 * sin2;
  Sp t4 = t2 * sin2;
  Sp Rp = Rs * (t3 - t4) / (t3 + t4);
  return (Rp + Rs) * 0.5f;
}
__device__ inline float fr_dielectric(float cosThetaI, float etaI, float etaT) {  // bsdf.cpp:149-170
  cosThetaI = clampf(cosThetaI, -1.0f, 1.0f);
  bool entering = cosThetaI > 0.f;
  if (!entering) { float t = etaI; etaI = etaT; etaT = t; cosThetaI = fabsf(cosThetaI); }
  float sinThetaI = sqrtf(mx(0.0f, 1 - cosThetaI * cosThetaI));
  float sinThetaT = etaI / etaT * sinThetaI;
  if (sinThetaT >= 1) return 1;
  float cosThetaT = sqrtf(mx(0.0f, 1 - sinThetaT * sinThetaT));
  float Rparl = ((etaT * cosThetaI) - (etaI * cosThetaT)) / ((etaT * cosThetaI) + (etaI * cosThetaT));
  float Rperp = ((etaI * cosThetaI) - (etaT * cosThetaT)) / ((etaI * cosThetaI) + (etaT * cosThetaT));
  return (Rparl * Rparl + Rperp * Rperp) / 2;
}
__device__ __forceinline__ Sp fresnel_eval(const Lobe &b, float cosI) {  // |cos| passed, bsdf.cpp:172-180
  if (b.fresnel == FR_CONDUCTOR) return fr_conductor(fabsf(cosI), sp(1.0f), b.etaT, b.k);
  if (b.fresnel == FR_NOOP) return sp(1.0f);  // bsdf.h:96-99
  return sp(fr_dielectric(fabsf(cosI), b.etaA, b.etaB));
}

__device__ __forceinline__ float cos2t(V3 w) { return w.z * w.z; }
__device__ __forceinline__ float sin2t(V3 w) { return mx(0.0f, 1.0f - cos2t(w)); }
__device__ __forceinline__ bool same_hemi(V3 a, V3 b) { return a.z * b.z > 0; }
__device__ inline float ggx_D(float alpha, V3 wh) {  // microfacet.cpp:9-18
  float tan2Th = sin2t(wh) / cos2t(wh);
  if (is_inf(tan2Th)) return 0;
  float cos4Th = cos2t(wh) * cos2t(wh);
  float alpha2 = alpha * alpha;
  float e = tan2Th / alpha2;
  return 1 / (kPi * alpha2 * cos4Th * (1 + e) * (1 + e));
}
__device__ inline float ggx_lambda(float alpha, V3 w) {  // microfacet.cpp:20-26
  float absTanTheta = fabsf(sqrtf(sin2t(w)) / w.z);
  if (is_inf(absTanTheta)) return 0.;
  float a2t2 = (alpha * absTanTheta) * (alpha * absTanTheta);
  return (-1 + sqrtf(1.f + a2t2)) / 2;
}
__device__ __forceinline__ float ggx_G(float alpha, V3 wi, V3 wo) {
  return 1 / (1 + ggx_lambda(alpha, wi) + ggx_lambda(alpha, wo));
}
__device__ __forceinline__ float ggx_pdf(float alpha, V3 wh) { return ggx_D(alpha, wh) * fabsf(wh.z); }
__device__ inline V3 ggx_sample_wh(float alpha, V3 wo, float u0, float u1) {  // microfacet.cpp:28-40
  float tanTheta2 = alpha * alpha * u0 / (1.0f - u0);
  float cosTheta = 1 / sqrtf(1 + tanTheta2);
  float sinTheta = sqrtf(mx(0.0f, 1.0f - cosTheta * cosTheta));
  float phi = (2 * kPi) * u1;
  V3 wh = mk(sinTheta * dcos(phi), sinTheta * dsin(phi), cosTheta);
  if (!same_hemi(wo, wh)) wh = neg(wh);
  return wh;
}
__device__ __forceinline__ V3 reflect(V3 wo, V3 n) { return add(neg(wo), mul(n, 2 * dot(wo, n))); }
__device__ __forceinline__ bool refract(V3 wi, V3 n, float eta, V3 *wt) {  // bsdf.h:105-115
  float cosThetaI = dot(n, wi);
  float sin2ThetaI = mx(0.0f, 1 - cosThetaI * cosThetaI);
  float sin2ThetaT = eta * eta * sin2ThetaI;
  if (sin2ThetaT >= 1) return false;
  float cosThetaT = sqrtf(1 - sin2ThetaT);
  *wt = add(mul(neg(wi), eta), mul(n, eta * cosThetaI - cosThetaT));
  return true;
}

// ------------------------------------------------------- lobes
__device__ inline Sp lobe_f(const Lobe &b, V3 wi, V3 wo) {
  if (b.kind == BX_LAMBERT) return b.R * kInvPi;  // lambert.cpp:5-8
  if (b.kind == BX_OREN_NAYAR) {                   // orennayar.cpp:12-37
    float sinThetaI = sqrtf(sin2t(wi)), sinThetaO = sqrtf(sin2t(wo));
    float cosTerm = 0;
    // 1e-4 is a double literal: the comparison is made in double
    if ((double)sinThetaI > 1e-4 && (double)sinThetaO > 1e-4) {
      // sampling.h:22-29 (the zero test there compares a function address,
      // so the division is always taken)
      float sinPhiI = clampf(wi.y / sinThetaI, -1.0f, 1.0f), cosPhiI = clampf(wi.x / sinThetaI, -1.0f, 1.0f);
      float sinPhiO = clampf(wo.y / sinThetaO, -1.0f, 1.0f), cosPhiO = clampf(wo.x / sinThetaO, -1.0f, 1.0f);
      float dCos = cosPhiI * cosPhiO + sinPhiI * sinPhiO;
      cosTerm = mx(0.0f, dCos);
    }
    float sinAlpha, tanBeta;
    if (fabsf(wi.z) > fabsf(wo.z)) {
      sinAlpha = sinThetaO;
      tanBeta = sinThetaI / fabsf(wi.z);
    } else {
      sinAlpha = sinThetaI;
      tanBeta = sinThetaO / fabsf(wo.z);
    }
    return b.R * kInvPi * (b.A + b.B * cosTerm * sinAlpha * tanBeta);
  }
  if (b.kind == BX_MF_REFL) {  // microfacet.cpp:55-66 (G(wi, wi) as in the reference)
    float cosThetaO = fabsf(wo.z), cosThetaI = fabsf(wi.z);
    if (cosThetaI == 0 || cosThetaO == 0) return sp(0);
    V3 wh = add(wi, wo);
    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return sp(0);
    wh = normalize(wh);
    Sp F = fresnel_eval(b, dot(wi, wh));
    return b.R * ggx_D(b.alpha, wh) * ggx_G(b.alpha, wi, wi) * F / (4 * cosThetaI * cosThetaO);
  }
  if (b.kind == BX_MF_TRANS) {  // microfacet.cpp:87-106
    if (same_hemi(wi, wo)) return sp(0);
    float cosThetaO = wo.z, cosThetaI = wi.z;
    if (cosThetaI == 0 || cosThetaO == 0) return sp(0);
    float eta = wo.z > 0 ? (b.etaB / b.etaA) : (b.etaA / b.etaB);
    V3 wh = normalize(add(wo, mul(wi, eta)));
    if (wh.z < 0) wh = neg(wh);
    Sp F = sp(fr_dielectric(fabsf(dot(wo, wh)), b.etaA, b.etaB));
    float sqrtDenom = dot(wo, wh) + eta * dot(wi, wh);
    float factor = b.photon ? 1 : (1 / eta);
    return (sp(1.f) - F) * b.T *
           fabsf(ggx_D(b.alpha, wh) * ggx_G(b.alpha, wi, wo) * eta * eta * absdot(wi, wh) * absdot(wo, wh) *
                 factor * factor / (cosThetaI * cosThetaO * sqrtDenom * sqrtDenom));
  }
  return sp(0);  // specular lobes (specular.h:31-34, 50-53, 72-75)
}
__device__ inline float lobe_pdf(const Lobe &b, V3 wi, V3 wo) {
  // Lambert and Oren-Nayar use BxDF::pdf (bsdf.cpp:117-121)
  if (b.kind == BX_LAMBERT || b.kind == BX_OREN_NAYAR) return same_hemi(wi, wo) ? fabsf(wi.z) * kInvPi : 0;
  if (b.kind == BX_MF_REFL) {
    if (!same_hemi(wi, wo)) return 0;
    V3 wh = normalize(add(wi, wo));
    return ggx_pdf(b.alpha, wh) / (4 * dot(wo, wh));
  }
  if (b.kind == BX_MF_TRANS) {
    if (same_hemi(wo, wi)) return 0;
    float eta = wo.z > 0 ? (b.etaB / b.etaA) : (b.etaA / b.etaB);
    V3 wh = normalize(add(wo, mul(wi, eta)));
    float sqrtDenom = dot(wo, wh) + eta * dot(wi, wh);
    float dwh_dwi = fabsf((eta * eta * dot(wi, wh)) / (sqrtDenom * sqrtDenom));
    return ggx_pdf(b.alpha, wh) * dwh_dwi;
  }
  return 0;
}
// bsdf.cpp:109-115 (cosine-weighted default), microfacet.cpp:42-53 / 76-85,
// specular.cpp:5-54.  *pdf is left untouched where the reference does not
// write it.
__device__ inline Sp lobe_sample(const Lobe &b, V3 wo, float u0, float u1, V3 *wi, float *pdf, int *stype) {
  if (b.kind == BX_LAMBERT || b.kind == BX_OREN_NAYAR) {
    *wi = cosine_hemisphere(u0, u1);
    if (wo.z < 0) wi->z *= -1;
    *pdf = lobe_pdf(b, *wi, wo);
    return lobe_f(b, *wi, wo);
  }
  if (b.kind == BX_MF_REFL) {
    if (wo.z == 0) return sp(0);
    V3 wh = ggx_sample_wh(b.alpha, wo, u0, u1);
    *wi = reflect(wo, wh);
    if (!same_hemi(*wi, wo)) return sp(0);
    *pdf = ggx_pdf(b.alpha, wh) / (4 * dot(wo, wh));
    return lobe_f(b, *wi, wo);
  }
  if (b.kind == BX_MF_TRANS) {  // *pdf is never written
    if (wo.z == 0) return sp(0);
    V3 wh = ggx_sample_wh(b.alpha, wo, u0, u1);
    float eta = wo.z > 0 ? (b.etaA / b.etaB) : (b.etaB / b.etaA);
    if (!refract(wo, wh, eta, wi)) return sp(0);
    return lobe_f(b, *wi, wo);
  }
  if (b.kind == BX_SPEC_REFL) {  // specular.cpp:19-24
    *wi = mk(-wo.x, -wo.y, wo.z);
    *pdf = 1;
    return fresnel_eval(b, wi->z) * b.R / fabsf(wi->z);
  }
  if (b.kind == BX_SPEC_TRANS) {  // specular.cpp:5-17
    bool entering = wo.z > 0;
    float etaI = entering ? b.etaA : b.etaB;
    float etaT = entering ? b.etaB : b.etaA;
    V3 n = mk(0, 0, 1);
    if (dot(n, wo) < 0.f) n = neg(n);  // faceforward
    if (!refract(wo, n, etaI / etaT, wi)) return sp(0);
    *pdf = 1.0f;
    Sp ft = b.T * (sp(1.0f) - sp(fr_dielectric(fabsf(wi->z), b.etaA, b.etaB)));
    if (!b.photon) mulby(ft, sp((etaI * etaI) / (etaT * etaT)));
    return ft / fabsf(wi->z);
  }
  // FresnelSpecular, specular.cpp:28-54
  float F = fr_dielectric(wo.z, b.etaA, b.etaB);
  if (u0 < F) {
    *wi = mk(-wo.x, -wo.y, wo.z);
    *stype = SPEC | REFL;
    *pdf = F;
    return b.R * F / fabsf(wi->z);
  }
  bool entering = wo.z > 0;
  float etaI = entering ? b.etaA : b.etaB;
  float etaT = entering ? b.etaB : b.etaA;
  V3 n = mk(0, 0, 1);
  if (dot(n, wo) < 0.f) n = neg(n);
  if (!refract(wo, n, etaI / etaT, wi)) return sp(0);
  Sp ft = b.T * (1 - F);
  if (!b.photon) mulby(ft, sp((etaI * etaI) / (etaT * etaT)));
  *stype = SPEC | TRANS;
  *pdf = 1 - F;
  return ft / fabsf(wi->z);
}

// ------------------------------------------------------- material -> lobes
// A BSDF keeps the material record's address and one 4-bit code per lobe (up
// to five, Uber's count); a lobe's parameters are re-derived from the record
// where they are used, so a shading lane holds one lobe in registers at a
// time, whatever the material.
enum {
  LC_LAMBERT = 0,    // clamp(rec[1..3])                        matte, plastic Kd
  LC_OREN_NAYAR,     // clamp(rec[1..3]), sigma rec[4]          matte, sigma != 0
  LC_METAL,          // GGX, conductor eta rec[1..3] k rec[4..6], alpha rec[8]
  LC_GLASS_SPEC,     // FresnelSpecular Kr rec[1..3] Kt rec[4..6], 1 | rec[9]
  LC_GLASS_MF_R,     // GGX reflection Kr, alpha rec[8], dielectric 1 | rec[9]
  LC_GLASS_MF_T,     // GGX transmission Kt, alpha rec[8], 1 | rec[9]
  LC_MIRROR,         // SpecularReflection clamp(rec[1..3]), FresnelNoOp
  LC_PLASTIC_MF,     // GGX clamp(rec[4..6]), alpha rec[8], dielectric 1.5 | 1
  LC_UBER_OPEN,      // SpecularTransmission clamp(1 - op), 1 | 1
  LC_UBER_DIFF,      // Lambert op * clamp(Kd)
  LC_UBER_MF,        // GGX op * clamp(Ks), alpha rec[17], dielectric 1 | e
  LC_UBER_SPEC_R,    // SpecularReflection op * clamp(Kr), dielectric 1 | e
  LC_UBER_SPEC_T,    // SpecularTransmission op * clamp(Kt), 1 | e
};
__device__ __forceinline__ int lc_type(int code) {
  switch (code) {
    case LC_LAMBERT: case LC_OREN_NAYAR: case LC_UBER_DIFF: return REFL | DIFF;
    case LC_METAL: case LC_GLASS_MF_R: case LC_PLASTIC_MF: case LC_UBER_MF: return REFL | GLOSSY;
    case LC_GLASS_SPEC: return REFL | TRANS | SPEC;
    case LC_GLASS_MF_T: return TRANS | GLOSSY;
    case LC_MIRROR: case LC_UBER_SPEC_R: return REFL | SPEC;
    default: return TRANS | SPEC;  // LC_UBER_OPEN, LC_UBER_SPEC_T
  }
}

struct Bsdf {
  V3 ng, ns, u, v;
  const float *m;  // material record
  uint32_t codes;  // lobe i: bits [4i, 4i + 4)
  uint32_t types;  // lobe i: BxDF type flags at bits [5i, 5i + 5)
  int n, photon;
  __device__ __forceinline__ V3 to_local(V3 w) const { return mk(dot(w, u), dot(w, v), dot(w, ns)); }
  __device__ __forceinline__ V3 to_world(V3 w) const {
    return mk(u.x * w.x + v.x * w.y + ns.x * w.z, u.y * w.x + v.y * w.y + ns.y * w.z,
              u.z * w.x + v.z * w.y + ns.z * w.z);
  }
  __device__ __forceinline__ int type_of(int i) const { return (int)(types >> (5 * i)) & 31; }
  __device__ __forceinline__ bool matches(int i, int t) const { int ty = type_of(i); return (ty & t) == ty; }
  __device__ __forceinline__ void push(int code) {
    codes |= (uint32_t)code << (4 * n);
    types |= (uint32_t)lc_type(code) << (5 * n);
    ++n;
  }
};

__device__ __forceinline__ Sp rec3(const float *m, int k) { return sp3(gld(m + k), gld(m + k + 1), gld(m + k + 2)); }
// Spectrum::clamp() (spectrum.h:89-93): [0, max float] whatever bounds are passed
__device__ __forceinline__ Sp clamp0(Sp s) {
  return sp3(clampf(s.r, 0, kFltMax), clampf(s.g, 0, kFltMax), clampf(s.b, 0, kFltMax));
}
__device__ __forceinline__ Sp uber_op(const float *m) { return rec3(m, 13); }

__device__ inline Lobe lobe_at(const Bsdf &bs, int i) {
  const float *m = bs.m;
  int code = (int)(bs.codes >> (4 * i)) & 15;
  Lobe b;
  b.kind = BX_LAMBERT; b.type = lc_type(code); b.fresnel = FR_DIELECTRIC; b.photon = bs.photon;
  b.R = sp(0); b.T = sp(0); b.etaT = sp(0); b.k = sp(0);
  b.alpha = 0; b.etaA = 1.0f; b.etaB = 1.0f; b.A = 0; b.B = 0;
  switch (code) {
    case LC_LAMBERT: b.R = clamp0(rec3(m, 1)); break;
    case LC_OREN_NAYAR: {  // orennayar.cpp:4-10
      b.kind = BX_OREN_NAYAR;
      b.R = clamp0(rec3(m, 1));
      float sigma = clampf(gld(m + 4), 0, kFltMax) * 0.01745329251994329576923690768489f;
      float sigma2 = sigma * sigma;
      b.A = 1.f - (sigma2 / (2.f * (sigma2 + 0.33f)));
      b.B = 0.45f * sigma2 / (sigma2 + 0.09f);
      break;
    }
    case LC_METAL:
      b.kind = BX_MF_REFL; b.R = sp(1.0f); b.alpha = gld(m + 8); b.fresnel = FR_CONDUCTOR;
      b.etaT = rec3(m, 1); b.k = rec3(m, 4);
      break;
    case LC_GLASS_SPEC:
      b.kind = BX_FRESNEL_SPEC; b.R = rec3(m, 1); b.T = rec3(m, 4); b.etaB = gld(m + 9);
      break;
    case LC_GLASS_MF_R:
      b.kind = BX_MF_REFL; b.R = rec3(m, 1); b.alpha = gld(m + 8); b.etaB = gld(m + 9);
      break;
    case LC_GLASS_MF_T:
      b.kind = BX_MF_TRANS; b.T = rec3(m, 4); b.alpha = gld(m + 8); b.etaB = gld(m + 9);
      break;
    case LC_MIRROR:
      b.kind = BX_SPEC_REFL; b.R = clamp0(rec3(m, 1)); b.fresnel = FR_NOOP;
      break;
    case LC_PLASTIC_MF:
      b.kind = BX_MF_REFL; b.R = clamp0(rec3(m, 4)); b.alpha = gld(m + 8); b.etaA = 1.5f; b.etaB = 1.0f;
      break;
    case LC_UBER_OPEN:
      b.kind = BX_SPEC_TRANS; b.T = clamp0(sp(1.0f) - uber_op(m));
      break;
    case LC_UBER_DIFF: b.R = uber_op(m) * clamp0(rec3(m, 1)); break;
    case LC_UBER_MF:
      b.kind = BX_MF_REFL; b.R = uber_op(m) * clamp0(rec3(m, 4)); b.alpha = gld(m + 17); b.etaB = gld(m + 18);
      break;
    case LC_UBER_SPEC_R:
      b.kind = BX_SPEC_REFL; b.R = uber_op(m) * clamp0(rec3(m, 7)); b.etaB = gld(m + 18);
      break;
    default:  // LC_UBER_SPEC_T
      b.kind = BX_SPEC_TRANS; b.T = uber_op(m) * clamp0(rec3(m, 10)); b.etaB = gld(m + 18);
      break;
  }
  return b;
}

// Material record -> lobe list, in the reference's addBxDF order
// (matte.cpp:9-21, metal.cpp:8-17, glass.cpp:9-37, mirror.cpp:7-14,
// plastic.cpp:8-24, uber.cpp:9-42); alpha comes precomputed in the record.
__device__ inline void make_bsdf(const float *m, V3 ng, V3 ns, V3 u, V3 v, bool photon, Bsdf *bs) {
  bs->ng = ng; bs->ns = ns; bs->u = u; bs->v = v;
  bs->m = m; bs->codes = 0; bs->types = 0; bs->n = 0; bs->photon = photon;
  int type = (int)gld(m + 0);
  if (type == 0) {
    if (!black(clamp0(rec3(m, 1)))) bs->push(clampf(gld(m + 4), 0, kFltMax) == 0 ? LC_LAMBERT : LC_OREN_NAYAR);
  } else if (type == 1) {
    bs->push(LC_METAL);
  } else if (type == 2) {
    Sp R = rec3(m, 1), T = rec3(m, 4);
    if (black(R) && black(T)) return;
    if (gld(m + 7) == 0) {
      bs->push(LC_GLASS_SPEC);
    } else {
      if (!black(R)) bs->push(LC_GLASS_MF_R);
      if (!black(T)) bs->push(LC_GLASS_MF_T);
    }
  } else if (type == 3) {
    if (!black(clamp0(rec3(m, 1)))) bs->push(LC_MIRROR);
  } else if (type == 4) {
    if (!black(clamp0(rec3(m, 1)))) bs->push(LC_LAMBERT);
    if (!black(clamp0(rec3(m, 4)))) bs->push(LC_PLASTIC_MF);
  } else {
    Sp op = uber_op(m);
    if (!black(clamp0(sp(1.0f) - op))) bs->push(LC_UBER_OPEN);
    if (!black(op * clamp0(rec3(m, 1)))) bs->push(LC_UBER_DIFF);
    if (!black(op * clamp0(rec3(m, 4)))) bs->push(LC_UBER_MF);
    if (!black(op * clamp0(rec3(m, 7)))) bs->push(LC_UBER_SPEC_R);
    if (!black(op * clamp0(rec3(m, 10)))) bs->push(LC_UBER_SPEC_T);
  }
}

// ------------------------------------------------------- BSDF (bsdf.cpp)
__device__ __forceinline__ int bsdf_num(const Bsdf &bs, int type) {  // bsdf.cpp:4-12
  int n = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) n += (i < bs.n && bs.matches(i, type)) ? 1 : 0;
  return n;
}
__device__ inline Sp bsdf_f(const Bsdf &bs, V3 wi, V3 wo, int type) {  // bsdf.cpp:17-29
  Sp r = sp(0);
  V3 swi = bs.to_local(wi), swo = bs.to_local(wo);
#pragma unroll 1
  for (int i = 0; i < bs.n; ++i)
    if (bs.matches(i, type)) acc(r, lobe_f(lobe_at(bs, i), swi, swo));
  return r;
}
__device__ inline float bsdf_pdf(const Bsdf &bs, V3 wi, V3 wo, int type) {  // bsdf.cpp:85-103
  if (bs.n == 0) return 0;
  V3 swi = bs.to_local(wi), swo = bs.to_local(wo);
  if (swo.z == 0) return 0;
  float pdf = 0;
  int m = 0;
#pragma unroll 1
  for (int i = 0; i < bs.n; ++i)
    if (bs.matches(i, type)) { ++m; pdf += lobe_pdf(lobe_at(bs, i), swi, swo); }
  return m > 0 ? pdf / m : 0.f;
}
// bsdf.cpp:31-83; *pdf and *stype stay untouched on the swo.z == 0 exit
__device__ inline Sp bsdf_sample(const Bsdf &bs, V3 *wi, V3 wo, float u0, float u1, float *pdf, int type, int *stype) {
  int matching = bsdf_num(bs, type);
  if (matching == 0) { *pdf = 0; *stype = 0; return sp(0); }
  int pos = min((int)floorf(u0 * matching), matching - 1);
  int count = pos, sel = 0;
#pragma unroll 1
  for (int i = 0; i < bs.n; ++i)
    if (bs.matches(i, type) && count-- == 0) { sel = i; break; }
  float nu0 = mn(u0 * matching - pos, kOneMinusEps);
  V3 swi = mk(0, 0, 0), swo = bs.to_local(wo);
  if (swo.z == 0) return sp(0);
  const int sel_type = bs.type_of(sel);
  *pdf = 0;
  *stype = sel_type;
  Sp f = lobe_sample(lobe_at(bs, sel), swo, nu0, u1, &swi, pdf, stype);
  if (*pdf == 0) { *stype = 0; return sp(0); }
  *wi = bs.to_world(swi);
  bool hasSpecular = (sel_type & SPEC) != 0;
  if (!hasSpecular && matching > 1) {
#pragma unroll 1
    for (int i = 0; i < bs.n; ++i)
      if (i != sel && bs.matches(i, type)) *pdf += lobe_pdf(lobe_at(bs, i), swi, swo);
  }
  if (matching > 1) *pdf /= matching;
  if (!hasSpecular && matching > 1) {
    bool refl = dot(*wi, bs.ng) * dot(wo, bs.ng) > 0;
    f = sp(0);
#pragma unroll 1
    for (int i = 0; i < bs.n; ++i) {
      int ty = bs.type_of(i);
      bool hr = (ty & REFL) != 0, ht = (ty & TRANS) != 0;
      if (bs.matches(i, type) && ((refl && hr) || (!refl && ht))) acc(f, lobe_f(lobe_at(bs, i), swi, swo));
    }
  }
  return f;
}

}  // namespace dev
}  // namespace orion_amd
