// Hero-wavelength spectral samples.
//
// Restates shaders/raytracing/spectralDistribution.hlsl:21-393 (wavelength set, sigmoid-polynomial
// reflectance spectra, D65-lit emission, Russian roulette, secondary-wavelength termination, XYZ
// projection) with the constants of shaders/sharedIncludes/spectralConstants.h:19-26.
//
// Colours reach the kernels already converted to sigmoid-polynomial coefficients (host side, through
// the RGB->coefficient table of SpectralUtility.cpp:253-438), so a hit evaluates sigmoid(poly(lambda))
// directly instead of sampling the 64^3 table; the arithmetic after the table lookup is unchanged.
#pragma once
#include "common.h"
#include "ymath.h"

namespace ypt {

static constexpr float LAMBDA_MIN = 380.f;        // CIE_LUT_LAMBDA_MIN
static constexpr float LAMBDA_MAX = 700.f;        // CIE_LUT_LAMBDA_MAX
static constexpr int CIE_RES = 321;               // CIE_LUT_RESOLUTION
static constexpr float CIE_Y_SUM_INV = 0.00936338721f;
static constexpr float CIE_D65_SUM_INV = 0.00003334595f;
static constexpr float LAMBDA_RANGE = 320.f;      // SPECTRAL_SAMPLE_LAMBDA_RANGE
static constexpr float LAMBDA_PDF = 1.f / 320.f;  // SPECTRAL_SAMPLE_LAMBDA_PDF

// Tables the spectral code reads: CIE xyz (float4 per nm, w unused) and D65, both 380..700 nm.
struct SpectralTables {
  const float* cie_xyz4;  // 321 * 4
  const float* d65;       // 321
};

// 1-D table lookup with the half-texel remap of sampleLUT (materialsCommonResources.hlsl:132-143)
// and linear filtering with clamp addressing: value at continuous index u * (n - 1).
YPT_HD float lut1_coord(float u, int n) {
  float scale = (float)(n - 1) / (float)n;
  float bias = 0.5f / (float)n;
  float c = u * scale + bias;
  return c * (float)n - 0.5f;
}
YPT_HD void lut1_index(float x, int n, int& i0, int& i1, float& f) {
  x = clampf(x, 0.f, (float)(n - 1));
  float fl = floorf(x);
  i0 = (int)fl;
  i1 = i0 + 1 < n ? i0 + 1 : n - 1;
  f = x - fl;
}

YPT_HD float sigmoid(float x) {  // spectralDistribution.hlsl:85-89
  if (is_inf(x)) return x > 0.f ? 1.f : 0.f;
  return 0.5f * x / sqrtf(1.0f + x * x) + 0.5f;
}
YPT_HD float sigmoid_inv(float x) {  // :91-96
  if (x == 1.0f) return as_float(0x7f800000u);
  if (x == 0.0f) return as_float(0xff800000u);
  return (x - 0.5f) / sqrtf(x * (1.f - x));
}

YPT_HD float hero_wavelength(float u) { return LAMBDA_MIN + u * LAMBDA_RANGE; }  // :43-47

struct Spectrum {
  float s[4];
  YPT_HD Spectrum() { s[0] = s[1] = s[2] = s[3] = 0.f; }
  YPT_HD explicit Spectrum(float v) { s[0] = s[1] = s[2] = s[3] = v; }
  YPT_HD Spectrum(float a, float b, float c, float d) { s[0] = a; s[1] = b; s[2] = c; s[3] = d; }
  YPT_HD void set(float v) { s[0] = s[1] = s[2] = s[3] = v; }
  YPT_HD float operator[](int i) const { return s[i]; }
  YPT_HD bool all_equal(float v) const { return s[0] == v && s[1] == v && s[2] == v && s[3] == v; }
  YPT_HD float max_value() const {  // getMaxSampleValue: v = max(s[i], v)
    float v = s[0];
    for (int i = 1; i < 4; ++i) v = fmax_(s[i], v);
    return v;
  }
  // applyRussianRoulette (:124-138): true terminates the path.
  YPT_HD bool russian_roulette(float rand, float normalization = 1.f) {
    float p = max_value() / normalization;
    p = clampf(p, 0.05f, 1.f);
    if (rand > p) return true;
    float inv = 1.f / p;
    for (int i = 0; i < 4; ++i) s[i] = s[i] * inv;
    return false;
  }
  YPT_HD void terminate_secondary() {  // :330-337
    s[0] *= 4.f;
    s[1] = 0.f; s[2] = 0.f; s[3] = 0.f;
  }
  YPT_HD void undo_secondary_rescale() { s[0] *= (1.f / 4.f); }
  YPT_HD bool has_nan() const { return is_nan(s[0]) || is_nan(s[1]) || is_nan(s[2]) || is_nan(s[3]); }
};

YPT_HD Spectrum pick(bool c, const Spectrum& a, const Spectrum& b) {
  Spectrum r;
  for (int i = 0; i < 4; ++i) {
    float x = a.s[i], y = b.s[i];
    r.s[i] = c ? x : y;
  }
  return r;
}

YPT_HD Spectrum operator*(const Spectrum& a, const Spectrum& b) { return Spectrum(a.s[0] * b.s[0], a.s[1] * b.s[1], a.s[2] * b.s[2], a.s[3] * b.s[3]); }
YPT_HD Spectrum operator/(const Spectrum& a, const Spectrum& b) { return Spectrum(a.s[0] / b.s[0], a.s[1] / b.s[1], a.s[2] / b.s[2], a.s[3] / b.s[3]); }
YPT_HD Spectrum operator+(const Spectrum& a, const Spectrum& b) { return Spectrum(a.s[0] + b.s[0], a.s[1] + b.s[1], a.s[2] + b.s[2], a.s[3] + b.s[3]); }
YPT_HD Spectrum operator-(const Spectrum& a, const Spectrum& b) { return Spectrum(a.s[0] - b.s[0], a.s[1] - b.s[1], a.s[2] - b.s[2], a.s[3] - b.s[3]); }
YPT_HD Spectrum operator*(const Spectrum& a, float v) { return Spectrum(a.s[0] * v, a.s[1] * v, a.s[2] * v, a.s[3] * v); }
YPT_HD Spectrum operator/(const Spectrum& a, float v) {  // one reciprocal, four products (as vec3 / float)
  const float r = 1.f / v;
  return Spectrum(a.s[0] * r, a.s[1] * r, a.s[2] * r, a.s[3] * r);
}
YPT_HD Spectrum operator+(const Spectrum& a, float v) { return Spectrum(a.s[0] + v, a.s[1] + v, a.s[2] + v, a.s[3] + v); }
YPT_HD Spectrum operator-(const Spectrum& a, float v) { return Spectrum(a.s[0] - v, a.s[1] - v, a.s[2] - v, a.s[3] - v); }

// getSpectralSampleLambdas (:21-41): hero wavelength and three more 80 nm apart, wrapped into range.
struct Lambdas {
  float l[4];
};
YPT_HD Lambdas spectral_lambdas(float hero) {
  Lambdas r;
  const float step = LAMBDA_RANGE / 4.f;
  float prev = hero;
  r.l[0] = prev;
  for (int i = 1; i < 4; ++i) {
    float lambda = prev + step;
    if (lambda > LAMBDA_MAX) lambda = LAMBDA_MIN + (lambda - LAMBDA_MAX);
    prev = lambda;
    r.l[i] = prev;
  }
  return r;
}

YPT_HD float lambda_to_unit(float lambda) { return (lambda - LAMBDA_MIN) / (LAMBDA_MAX - LAMBDA_MIN); }

YPT_HD float d65_at(const SpectralTables& t, float lambda) {  // getIlluminantCoeffForWavelength
  int i0, i1; float f;
  lut1_index(lut1_coord(lambda_to_unit(lambda), CIE_RES), CIE_RES, i0, i1, f);
  float a = t.d65[i0], b = t.d65[i1];
  return a + f * (b - a);
}
YPT_HD vec3 cie_xyz_at(const SpectralTables& t, float lambda) {  // getXYZCoeffsForWavelength
  int i0, i1; float f;
  lut1_index(lut1_coord(lambda_to_unit(lambda), CIE_RES), CIE_RES, i0, i1, f);
  vec3 a(t.cie_xyz4[i0 * 4 + 0], t.cie_xyz4[i0 * 4 + 1], t.cie_xyz4[i0 * 4 + 2]);
  vec3 b(t.cie_xyz4[i1 * 4 + 0], t.cie_xyz4[i1 * 4 + 1], t.cie_xyz4[i1 * 4 + 2]);
  return a + f * (b - a);
}

// setWithPolynomialCoeffs (:199-211)
YPT_HD Spectrum spectrum_from_coeffs(const float c[3], const Lambdas& L) {
  Spectrum r;
  for (int i = 0; i < 4; ++i) {
    float lambda = lambda_to_unit(L.l[i]);
    float polynom = c[0] * lambda * lambda + c[1] * lambda + c[2];
    r.s[i] = sigmoid(polynom);
  }
  return r;
}

// A colour converted on the host: sigmoid-polynomial coefficients of the (normalised) colour and the
// scale of the unbounded forms (setFromRGBUnbounded / setFromRGBUnboundedNoIlluminant, :140-180).
struct SpectralColor {
  float c[3];
  float scale;
};

// RGB -> sigmoid-polynomial coefficients via the 64^3 table (trilinear; grey colours exactly)
YPT_HD void rgb_coeffs_table(const float* table, vec3 v, float c[3]) {
  if (v.x == v.y && v.x == v.z) { c[0] = 0.f; c[1] = 0.f; c[2] = sigmoid_inv(v.x); return; }
  if (!table) { c[0] = 0.f; c[1] = 0.f; c[2] = as_float(0x7fc00000u); return; }  // table missing: NaN marks the sample
  const int R = 64;
  int i0[3], i1[3]; float f[3];
  for (int a = 0; a < 3; ++a) lut1_index(lut1_coord(v[a], R), R, i0[a], i1[a], f[a]);
  for (int k = 0; k < 3; ++k) {
    float cz[2];
    for (int zz = 0; zz < 2; ++zz) {
      int z = zz ? i1[2] : i0[2];
      float cy[2];
      for (int yy = 0; yy < 2; ++yy) {
        int y = yy ? i1[1] : i0[1];
        float a = table[((z * R + y) * R + i0[0]) * 4 + k];
        float b = table[((z * R + y) * R + i1[0]) * 4 + k];
        cy[yy] = a + f[0] * (b - a);
      }
      cz[zz] = cy[0] + f[1] * (cy[1] - cy[0]);
    }
    c[k] = cz[0] + f[2] * (cz[1] - cz[0]);
  }
}

YPT_HD Spectrum spectrum_bounded(const SpectralColor& col, const Lambdas& L) { return spectrum_from_coeffs(col.c, L); }
YPT_HD Spectrum spectrum_unbounded_no_illum(const SpectralColor& col, const Lambdas& L) {
  Spectrum r = spectrum_from_coeffs(col.c, L);
  return r * col.scale;
}
YPT_HD Spectrum spectrum_unbounded(const SpectralColor& col, const Lambdas& L, const SpectralTables& t) {
  Spectrum r = spectrum_from_coeffs(col.c, L);
  r = r * col.scale;
  for (int i = 0; i < 4; ++i) r.s[i] *= d65_at(t, L.l[i]) * CIE_D65_SUM_INV * 100.f;
  return r;
}

// ToXYZ (:345-365)
YPT_HD vec3 spectrum_to_xyz(const Spectrum& s, const Lambdas& L, const SpectralTables& t) {
  vec3 xyz(0.f);
  for (int i = 0; i < 4; ++i) {
    vec3 w = cie_xyz_at(t, L.l[i]);
    xyz = xyz + s.s[i] * w * safe_div(CIE_Y_SUM_INV, LAMBDA_PDF);
  }
  return xyz / 4.f;
}

}  // namespace ypt
