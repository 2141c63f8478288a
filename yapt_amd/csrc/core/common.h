// Shared scalar/vector vocabulary of the path-tracing core.
//
// Every function here is compiled twice: by hipcc for gfx950 device code and by g++ for the host
// (scene preparation, LUT integration, and the CPU checker). Both builds use -ffp-contract=off and
// IEEE-correct division and square root, so a function evaluated on the same inputs returns the
// same bits on both sides. Operation order follows the HLSL the functions restate.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define YPT_HD __host__ __device__ __forceinline__
#define YPT_DEVICE_CODE 1
#else
#include <cmath>
#include <cstdint>
#include <cstring>
#define YPT_HD inline
#endif

#include <stdint.h>

namespace ypt {

static constexpr float PI = 3.14159265f;        // shaders/common/commonMath.hlsl:5
static constexpr float INVPI = 0.31830988f;     // :6
static constexpr float PI_OVER_2 = 1.57079633f; // :7
static constexpr float PI_OVER_4 = 0.785398163f;// :8

YPT_HD float as_float(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
YPT_HD uint32_t as_uint(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
YPT_HD int32_t as_int(float f) { return (int32_t)as_uint(f); }
YPT_HD float as_float_i(int32_t i) { return as_float((uint32_t)i); }

// Element i of a small fixed-size array through unrolled value selects: a dynamic index into a
// register array (or a select between element addresses) would put the whole array in scratch
// memory. Values are copied unchanged. i must be below N. Types other than float and uint32_t
// provide their own pick() next to their definition.
YPT_HD float pick(bool c, float a, float b) { return c ? a : b; }
YPT_HD uint32_t pick(bool c, uint32_t a, uint32_t b) { return c ? a : b; }

template <class T, int N>
YPT_HD T sel_get(const T (&a)[N], uint32_t i) {
  T r = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = pick((uint32_t)k == i, a[k], r);
  return r;
}
template <class T, int N>
YPT_HD void sel_set(T (&a)[N], uint32_t i, const T& v) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = pick((uint32_t)k == i, v, a[k]);
}

// f32tof16 / f16tof32 with round-to-nearest-even (packed light-vertex normals, packed occlusion and
// denoiser features are kept at this precision)
YPT_HD uint32_t f32_to_f16(float f) {
  uint32_t x = as_uint(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t e = (x >> 23) & 0xFFu;
  uint32_t m = x & 0x7FFFFFu;
  if (e == 0xFFu) return sign | 0x7C00u | (m ? 0x200u : 0u);
  int ex = (int)e - 127 + 15;
  if (ex >= 31) return sign | 0x7C00u;
  if (ex <= 0) {
    if (ex < -10) return sign;
    m |= 0x800000u;
    uint32_t shift = (uint32_t)(14 - ex);
    uint32_t hm = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (hm & 1u))) hm++;
    return sign | hm;
  }
  uint32_t h = sign | ((uint32_t)ex << 10) | (m >> 13);
  uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return h;
}
YPT_HD float f16_to_f32(uint32_t h) {
  uint32_t sign = (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
  if (e == 0) {
    float v = (float)m * 5.9604644775390625e-8f;  // m * 2^-24, exact
    return sign ? -v : v;
  }
  if (e == 31) return as_float(sign | 0x7F800000u | (m << 13));
  return as_float(sign | ((e + 112u) << 23) | (m << 13));
}

YPT_HD float fmin_(float a, float b) { return a < b ? a : b; }   // HLSL min on non-NaN operands
YPT_HD float fmax_(float a, float b) { return a > b ? a : b; }
YPT_HD float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
YPT_HD float saturate(float x) { return clampf(x, 0.f, 1.f); }
YPT_HD float sqr(float x) { return x * x; }
YPT_HD float lerpf(float a, float b, float t) { return a + t * (b - a); }
YPT_HD float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
YPT_HD bool is_nan(float x) { return x != x; }
YPT_HD bool is_inf(float x) { return (as_uint(x) & 0x7fffffffu) == 0x7f800000u; }
YPT_HD bool is_finite(float x) { return (as_uint(x) & 0x7f800000u) != 0x7f800000u; }
YPT_HD float safe_div(float n, float d) { return d == 0.f ? 0.f : n / d; }   // common.hlsl:13-22
YPT_HD float safe_divisor(float x) { return x == 0.f ? 1.f : x; }             // common.hlsl:4

struct float2_ { float x, y; };

struct vec2 {
  float x, y;
  YPT_HD vec2() : x(0.f), y(0.f) {}
  YPT_HD vec2(float a, float b) : x(a), y(b) {}
};

struct vec3 {
  float x, y, z;
  YPT_HD vec3() : x(0.f), y(0.f), z(0.f) {}
  YPT_HD explicit vec3(float s) : x(s), y(s), z(s) {}
  YPT_HD vec3(float a, float b, float c) : x(a), y(b), z(c) {}
  YPT_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  YPT_HD float& at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

YPT_HD vec3 operator+(vec3 a, vec3 b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
YPT_HD vec3 operator-(vec3 a, vec3 b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
YPT_HD vec3 operator*(vec3 a, vec3 b) { return vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
YPT_HD vec3 operator/(vec3 a, vec3 b) { return vec3(a.x / b.x, a.y / b.y, a.z / b.z); }
YPT_HD vec3 operator*(vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
YPT_HD vec3 operator*(float s, vec3 a) { return vec3(s * a.x, s * a.y, s * a.z); }
// one correctly rounded reciprocal and three products: within an ulp of three quotients, at a third
// of their cost on the GPU (both the device code and the CPU checker compile this definition)
YPT_HD vec3 operator/(vec3 a, float s) {
  const float r = 1.f / s;
  return vec3(a.x * r, a.y * r, a.z * r);
}
YPT_HD vec3 operator-(vec3 a) { return vec3(-a.x, -a.y, -a.z); }
YPT_HD float dot(vec3 a, vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
YPT_HD vec3 cross(vec3 a, vec3 b) {
  return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
YPT_HD float length(vec3 a) { return sqrtf(dot(a, a)); }
YPT_HD vec3 normalize(vec3 a) { return a / length(a); }
YPT_HD vec3 vabs(vec3 a) { return vec3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
YPT_HD vec3 vmin(vec3 a, vec3 b) { return vec3(fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)); }
YPT_HD vec3 vmax(vec3 a, vec3 b) { return vec3(fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)); }
YPT_HD bool is_zero(vec3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
YPT_HD float max3(vec3 a) { return fmax_(fmax_(a.x, a.y), a.z); }
YPT_HD vec3 reflect(vec3 i, vec3 n) { return i - (2.f * dot(n, i)) * n; }   // HLSL reflect
// HLSL refract(i, n, eta): zero vector on total internal reflection.
YPT_HD vec3 refract(vec3 i, vec3 n, float eta) {
  float ni = dot(n, i);
  float k = 1.f - eta * eta * (1.f - ni * ni);
  if (k < 0.f) return vec3(0.f);
  return eta * i - (eta * ni + sqrtf(k)) * n;
}

// Row-major 3x3: rows r0, r1, r2 (HLSL float3x3(r0, r1, r2)).
struct mat3 {
  vec3 r0, r1, r2;
  YPT_HD mat3() {}
  YPT_HD mat3(vec3 a, vec3 b, vec3 c) : r0(a), r1(b), r2(c) {}
};
YPT_HD vec3 mul(const mat3& m, vec3 v) { return vec3(dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)); }
// row vector times matrix: v.x * r0 + v.y * r1 + v.z * r2
YPT_HD vec3 mul(vec3 v, const mat3& m) { return (v.x * m.r0 + v.y * m.r1) + v.z * m.r2; }

// Row-major 4x4 (HLSL float4x4 as uploaded from glm column-major with the HLSL default packing).
struct mat4 {
  float m[16];  // m[row * 4 + col]
};
YPT_HD void mul4(const mat4& M, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r)
    out[r] = ((M.m[r * 4 + 0] * v[0] + M.m[r * 4 + 1] * v[1]) + M.m[r * 4 + 2] * v[2]) + M.m[r * 4 + 3] * v[3];
}

}  // namespace ypt
