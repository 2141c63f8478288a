// FP32 transcendental functions shared by the device kernels and the host build.
//
// Device libraries and the host libm round transcendentals differently, which would break bit-exact
// agreement of sampled directions between the GPU and the CPU checker. These routines use only IEEE
// +,-,*,/ and exact operations (floor, bit manipulation), so they return the same bits on both.
// Method and coefficients: Cephes single-precision library (S. L. Moshier), sinf.c / cosf.c
// (three-part Cody-Waite reduction by pi/4, degree-7/8 minimax polynomials), expf.c (reduction by
// ln2, degree-6 polynomial), logf.c (mantissa in [sqrt(1/2), sqrt(2)), degree-9 polynomial),
// atanf.c (reduction by tan(3pi/8) and tan(pi/8), odd degree-9 polynomial), atan2f.c (quadrant
// offset by pi), asinf.c (degree-11 odd polynomial, half-angle identity above 1/2) and acosf.c.
// Accuracy is a few ulp over the argument ranges the renderer uses (|x| < 8192 for sin/cos).
#pragma once
#include "common.h"

namespace ypt {

static constexpr float YM_FOPI = 1.27323954473516f;  // 4/pi
static constexpr float YM_DP1 = 0.78515625f;
static constexpr float YM_DP2 = 2.4187564849853515625e-4f;
static constexpr float YM_DP3 = 3.77489497744594108e-8f;

YPT_HD float ym_sin_poly(float x, float z) {
  return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
YPT_HD float ym_cos_poly(float z) {
  return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z - 0.5f * z + 1.0f;
}

YPT_HD void ysincos(float xin, float* s_out, float* c_out) {
  if (!is_finite(xin)) { float n = xin - xin; *s_out = n; *c_out = n; return; }
  float x = fabsf(xin);
  bool neg = xin < 0.f;
  uint32_t j = (uint32_t)(x * YM_FOPI);
  float y = (float)j;
  if (j & 1u) { j += 1u; y += 1.0f; }
  j &= 7u;
  int sign_s = neg ? -1 : 1;
  int sign_c = 1;
  if (j > 3u) { sign_s = -sign_s; sign_c = -sign_c; j -= 4u; }
  if (j > 1u) sign_c = -sign_c;
  x = ((x - y * YM_DP1) - y * YM_DP2) - y * YM_DP3;
  float z = x * x;
  float ps = ym_sin_poly(x, z);
  float pc = ym_cos_poly(z);
  float s, c;
  if (j == 1u || j == 2u) { s = pc; c = ps; } else { s = ps; c = pc; }
  *s_out = sign_s < 0 ? -s : s;
  *c_out = sign_c < 0 ? -c : c;
}
YPT_HD float ysin(float x) { float s, c; ysincos(x, &s, &c); return s; }
YPT_HD float ycos(float x) { float s, c; ysincos(x, &s, &c); return c; }

// x * 2^n for n in [-300, 300] through exact power-of-two multiplications.
YPT_HD float yldexp(float x, int n) {
  if (n > 127) { x *= as_float(0x7f000000u); n -= 127; if (n > 127) { x *= as_float(0x7f000000u); n -= 127; if (n > 127) n = 127; } }
  if (n < -126) { x *= as_float(0x00800000u); n += 126; if (n < -126) { x *= as_float(0x00800000u); n += 126; if (n < -126) n = -126; } }
  return x * as_float((uint32_t)(n + 127) << 23);
}

YPT_HD float yexp(float x) {
  if (x != x) return x;
  if (x > 88.72283905206835f) return as_float(0x7f800000u);
  if (x < -103.278929903431851103f) return 0.f;
  float z = floorf(1.44269504088896341f * x + 0.5f);
  x = (x - z * 0.693359375f) - z * -2.12194440e-4f;
  int n = (int)z;
  float x2 = x * x;
  float p = (((((1.9875691500e-4f * x + 1.3981999507e-3f) * x + 8.3334519073e-3f) * x + 4.1665795894e-2f) * x +
              1.6666665459e-1f) * x + 5.0000001201e-1f) * x2 + x + 1.0f;
  return yldexp(p, n);
}

YPT_HD float ylog(float x) {
  if (x != x) return x;
  if (x < 0.f) return as_float(0x7fc00000u);
  if (x == 0.f) return as_float(0xff800000u);
  if (is_inf(x)) return x;
  int eadj = 0;
  if (x < 1.17549435e-38f) { x *= 8388608.0f; eadj = -23; }  // subnormal: scale by 2^23
  uint32_t b = as_uint(x);
  int e = (int)((b >> 23) & 0xffu) - 126 + eadj;
  float m = as_float((b & 0x807fffffu) | 0x3f000000u);  // [0.5, 1)
  if (m < 0.707106781186547524f) { e -= 1; m = m + m - 1.0f; } else { m = m - 1.0f; }
  float z = m * m;
  float y = ((((((((7.0376836292e-2f * m - 1.1514610310e-1f) * m + 1.1676998740e-1f) * m - 1.2420140846e-1f) * m +
                 1.4249322787e-1f) * m - 1.6668057665e-1f) * m + 2.0000714765e-1f) * m - 2.4999993993e-1f) * m +
             3.3333331174e-1f) * m * z;
  float fe = (float)e;
  y += -2.12194440e-4f * fe;
  y += -0.5f * z;
  float r = m + y;
  r += 0.693359375f * fe;
  return r;
}

// HLSL pow for the non-negative bases the renderer uses.
YPT_HD float ypow(float x, float y) {
  if (y == 0.f || x == 1.f) return 1.f;
  if (x != x || y != y) return as_float(0x7fc00000u);
  if (x == 0.f) return y > 0.f ? 0.f : as_float(0x7f800000u);
  if (x < 0.f) return as_float(0x7fc00000u);
  return yexp(y * ylog(x));
}

// Cephes atanf: |x| above tan(3pi/8) maps to -1/|x| with offset pi/2, above tan(pi/8) to
// (|x| - 1) / (|x| + 1) with offset pi/4.
YPT_HD float yatan(float xin) {
  float x = fabsf(xin), y;
  if (x > 2.414213562373095f) {
    y = PI_OVER_2;
    x = -(1.f / x);
  } else if (x > 0.4142135623730950f) {
    y = PI_OVER_4;
    x = (x - 1.f) / (x + 1.f);
  } else {
    y = 0.f;
  }
  float z = x * x;
  y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z * x + x);
  return xin < 0.f ? -y : y;
}

// Cephes atan2f: the angle of (x, y) in [-pi, pi]
YPT_HD float yatan2(float y, float x) {
  if (x == 0.f) return y < 0.f ? -PI_OVER_2 : (y == 0.f ? 0.f : PI_OVER_2);
  if (y == 0.f) return x < 0.f ? PI : 0.f;
  float w = x < 0.f ? (y < 0.f ? -PI : PI) : 0.f;
  return w + yatan(y / x);
}

// Cephes asinf for |x| <= 1 (0 outside, as Cephes's domain error returns)
YPT_HD float yasin(float xin) {
  float a = fabsf(xin);
  if (a > 1.f) return 0.f;
  float z;
  if (a < 1e-4f) {
    z = a;
  } else {
    bool flag = a > 0.5f;
    float x;
    if (flag) {
      z = 0.5f * (1.f - a);
      x = sqrtf(z);
    } else {
      x = a;
      z = x * x;
    }
    z = (((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z + 7.4953002686e-2f) * z + 1.6666752422e-1f) * z * x) + x;
    if (flag) {
      z = z + z;
      z = PI_OVER_2 - z;
    }
  }
  return xin < 0.f ? -z : z;
}

// Cephes acosf for |x| <= 1 (0 outside)
YPT_HD float yacos(float x) {
  if (x < -1.f || x > 1.f) return 0.f;
  if (x < -0.5f) return PI - 2.f * yasin(sqrtf(0.5f * (1.f + x)));
  if (x > 0.5f) return 2.f * yasin(sqrtf(0.5f * (1.f - x)));
  return PI_OVER_2 - yasin(x);
}

}  // namespace ypt
