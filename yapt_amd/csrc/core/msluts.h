// Directional-albedo tables for microfacet energy compensation and their lookups.
//
// Table sizes and axes follow MultipleScatterLUTUtility.h:8-11 and MultipleScatterLUT.cpp:22-60;
// lookups follow multiScatter.hlsl:12-69. The reference samples them as textures with hardware
// linear filtering after the half-texel remap of sampleLUT (materialsCommonResources.hlsl:132-169);
// here the filter runs in FP32 with the same remap, identically on host and device.
#pragma once
#include "common.h"
#include "spectral.h"

namespace ypt {

static constexpr int SS_DIM = 32;     // SS_ALBEDO_LUT_DIM
static constexpr int SSMS_DIM = 16;   // SSMS_ALBEDO_LUT_DIM
static constexpr int TR_DIM = 16;     // SS_ALBEDO_TRANSLUCENT_LUT_DIM
static constexpr int SHEEN_DIM = 16;  // SS_ALBEDO_SHEEN_LUT_DIM

// Offsets (in floats) of each table inside one contiguous block.
static constexpr int MSL_SS_DIR = 0;                                   // [rough][cos] 32x32
static constexpr int MSL_SS_AVG = MSL_SS_DIR + SS_DIM * SS_DIM;        // [rough] 32
static constexpr int MSL_SSMS_DIR = MSL_SS_AVG + SS_DIM;               // [eta][rough][cos] 16^3
static constexpr int MSL_SSMS_AVG = MSL_SSMS_DIR + SSMS_DIM * SSMS_DIM * SSMS_DIM;  // [eta][rough]
static constexpr int MSL_TR_DIR_DENSER = MSL_SSMS_AVG + SSMS_DIM * SSMS_DIM;       // [eta][rough][cos]
static constexpr int MSL_TR_DIR_LIGHTER = MSL_TR_DIR_DENSER + TR_DIM * TR_DIM * TR_DIM;
static constexpr int MSL_TR_AVG_DENSER = MSL_TR_DIR_LIGHTER + TR_DIM * TR_DIM * TR_DIM;  // [eta][rough]
static constexpr int MSL_TR_AVG_LIGHTER = MSL_TR_AVG_DENSER + TR_DIM * TR_DIM;
static constexpr int MSL_SHEEN_DIR = MSL_TR_AVG_LIGHTER + TR_DIM * TR_DIM;           // [rough][cos]
static constexpr int MSL_TOTAL = MSL_SHEEN_DIR + SHEEN_DIM * SHEEN_DIM;

YPT_HD float lut1(const float* t, int n, float u) {
  int i0, i1; float f;
  lut1_index(lut1_coord(u, n), n, i0, i1, f);
  float a = t[i0], b = t[i1];
  return a + f * (b - a);
}
// t[y * w + x]
YPT_HD float lut2(const float* t, int w, int h, float u, float v) {
  int x0, x1, y0, y1; float fx, fy;
  lut1_index(lut1_coord(u, w), w, x0, x1, fx);
  lut1_index(lut1_coord(v, h), h, y0, y1, fy);
  float a = t[y0 * w + x0], b = t[y0 * w + x1], c = t[y1 * w + x0], d = t[y1 * w + x1];
  float ab = a + fx * (b - a);
  float cd = c + fx * (d - c);
  return ab + fy * (cd - ab);
}
// t[(z * h + y) * w + x]
YPT_HD float lut3(const float* t, int w, int h, int dd, float u, float v, float s) {
  int z0, z1; float fz;
  lut1_index(lut1_coord(s, dd), dd, z0, z1, fz);
  float a = lut2(t + z0 * w * h, w, h, u, v);
  float b = lut2(t + z1 * w * h, w, h, u, v);
  return a + fz * (b - a);
}

struct MSLuts {
  const float* base;  // MSL_TOTAL floats

  YPT_HD float ss_dir(float cosTheta, float lr) const { return lut2(base + MSL_SS_DIR, SS_DIM, SS_DIM, cosTheta, lr); }
  YPT_HD float ss_avg(float lr) const { return lut1(base + MSL_SS_AVG, SS_DIM, lr); }
  YPT_HD float ssms_dir(float eta, float cosTheta, float lr) const {
    float e = saturate((eta - 1.f) * 0.333333f);
    return lut3(base + MSL_SSMS_DIR, SSMS_DIM, SSMS_DIM, SSMS_DIM, cosTheta, lr, e);
  }
  YPT_HD float ssms_avg(float eta, float lr) const {
    float e = saturate((eta - 1.f) * 0.333333f);
    return lut2(base + MSL_SSMS_AVG, SSMS_DIM, SSMS_DIM, lr, e);
  }
  YPT_HD float tr_dir_lighter(float eta, float cosTheta, float lr) const {
    float e = saturate((eta - 0.3f) * (1.f / 0.69f));
    return lut3(base + MSL_TR_DIR_LIGHTER, TR_DIM, TR_DIM, TR_DIM, cosTheta, lr, e);
  }
  YPT_HD float tr_dir_denser(float eta, float cosTheta, float lr) const {
    float e = saturate((eta - 1.f) * 0.333333f);
    return lut3(base + MSL_TR_DIR_DENSER, TR_DIM, TR_DIM, TR_DIM, cosTheta, lr, e);
  }
  YPT_HD float tr_avg_lighter(float eta, float lr) const {
    float e = saturate((eta - 0.3f) * (1.f / 0.69f));
    return lut2(base + MSL_TR_AVG_LIGHTER, TR_DIM, TR_DIM, lr, e);
  }
  YPT_HD float tr_avg_denser(float eta, float lr) const {
    float e = saturate((eta - 1.f) * 0.333333f);
    return lut2(base + MSL_TR_AVG_DENSER, TR_DIM, TR_DIM, lr, e);
  }
  YPT_HD float sheen_dir(float cosTheta, float roughness) const {
    roughness = saturate(roughness - 0.07f);
    return lut2(base + MSL_SHEEN_DIR, SHEEN_DIM, SHEEN_DIM, cosTheta, roughness);
  }
};

}  // namespace ypt
