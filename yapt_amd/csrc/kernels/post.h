// The steps after the film on gfx950: feature prefilter and edge-avoiding a-trous denoiser
// (prefilterFeatures.comp, denoise.comp, denoiseFeatures.hlsl), XYZ -> linear sRGB with the gamut
// clamp, exposure and the ACES fitted operator (tonemap.frag, tonemapUtils.hlsl), and the 8-bit sRGB
// encoding of an RGBA8_SRGB render target (the FinalColor readback).
//
// One thread per pixel, 8x8 workgroups as in the reference. Every pass reads its neighbourhood from
// the previous pass's buffers; the a-trous footprint (5x5 taps at stride 2^pass) and the 3x3 variance
// prefilter are gathers, so these kernels are bound by L2/HBM gathers, not arithmetic.
#pragma once
#include "../core/bdpt.h"  // f32_to_f16 / f16_to_f32 (round to nearest even)

namespace ypt {

static constexpr float FEATURE_HALF_MAX = 65000.f;

// DenoiseFeatures (denoiseFeatures.hlsl:46-61)
struct DenoiseFeatures {
  vec3 albedo, normal;
  float depth, roughness, metalness, transparency, anisotropy, ior;
  float varAlbedo, varDepth, varMaterial;
};

struct DenoiseWeights {  // FeatureWeightParams
  float normalSigma, materialSigma, depthSigmaScale, varianceScale;
};

struct PostParams {
  mat4 uvToView, viewToWorld;
  uint32_t width, height;
  float depthSigmaScale, normalSigma, materialSigma, colorSigma, featureVarianceScale;
  uint32_t prefilter;
};

// featuresFromMoments (denoiseFeatures.hlsl:63-88)
__device__ __forceinline__ DenoiseFeatures features_from_moments(float4 l0, float4 l1, float4 l2, float4 l3) {
  DenoiseFeatures f;
  f.albedo = vec3(l0.x, l0.y, l0.z);
  f.varAlbedo = fmax_(l0.w - dot(f.albedo, f.albedo), 0.f);
  f.normal = vec3(l1.x, l1.y, l1.z);
  float coverage = l1.w;
  f.roughness = l2.x; f.metalness = l2.y; f.transparency = l2.z; f.anisotropy = l2.w;
  f.ior = l3.x;
  f.varMaterial = fmax_(l3.y - (l2.x * l2.x + l2.y * l2.y + l2.z * l2.z + l2.w * l2.w) - f.ior * f.ior, 0.f);
  if (coverage > 0.f) {
    f.depth = l3.z / coverage;
    f.varDepth = fmax_(l3.w / coverage - f.depth * f.depth, 0.f);
  } else {
    f.depth = 0.f;
    f.varDepth = 0.f;
  }
  return f;
}

__device__ __forceinline__ DenoiseFeatures load_moments(const float4* acc, size_t npix, uint32_t p) {
  return features_from_moments(acc[p], acc[npix + p], acc[2 * npix + p], acc[3 * npix + p]);
}

__device__ __forceinline__ uint32_t pack_half2(float a, float b) { return f32_to_f16(a) | (f32_to_f16(b) << 16); }

// writeDenoiseFeatures / readDenoiseFeatures (denoiseFeatures.hlsl:105-146): half precision except the depth
__device__ __forceinline__ void write_features(uint4* o0, uint4* o1, uint32_t p, const DenoiseFeatures& f) {
  float relVarDepth = f.depth > 0.f ? fmin_(f.varDepth / (f.depth * f.depth), FEATURE_HALF_MAX) : 0.f;
  o0[p] = make_uint4(pack_half2(f.albedo.x, f.albedo.y), pack_half2(f.albedo.z, relVarDepth), pack_half2(f.normal.x, f.normal.y),
                     pack_half2(f.normal.z, f.roughness));
  o1[p] = make_uint4(__float_as_uint(f.depth), pack_half2(f.metalness, f.anisotropy), pack_half2(f.transparency, f.ior),
                     pack_half2(fmin_(f.varAlbedo, FEATURE_HALF_MAX), fmin_(f.varMaterial, FEATURE_HALF_MAX)));
}

__device__ __forceinline__ DenoiseFeatures read_features(const uint4* i0, const uint4* i1, uint32_t p) {
  uint4 v0 = i0[p], v1 = i1[p];
  DenoiseFeatures f;
  f.albedo = vec3(f16_to_f32(v0.x & 0xFFFFu), f16_to_f32(v0.x >> 16), f16_to_f32(v0.y & 0xFFFFu));
  float relVarDepth = f16_to_f32(v0.y >> 16);
  f.normal = vec3(f16_to_f32(v0.z & 0xFFFFu), f16_to_f32(v0.z >> 16), f16_to_f32(v0.w & 0xFFFFu));
  f.roughness = f16_to_f32(v0.w >> 16);
  f.depth = __uint_as_float(v1.x);
  f.metalness = f16_to_f32(v1.y & 0xFFFFu);
  f.anisotropy = f16_to_f32(v1.y >> 16);
  f.transparency = f16_to_f32(v1.z & 0xFFFFu);
  f.ior = f16_to_f32(v1.z >> 16);
  f.varAlbedo = f16_to_f32(v1.w & 0xFFFFu);
  f.varMaterial = f16_to_f32(v1.w >> 16);
  f.varDepth = relVarDepth * f.depth * f.depth;
  return f;
}

// calculateVarianceNormalizedWeight (denoiseFeatures.hlsl:157-160)
__device__ __forceinline__ float variance_weight(float distSqr, float sigma, float varianceSum, float varianceScale) {
  return expf(-distSqr / (sigma * sigma + varianceScale * varianceSum + 1e-8f));
}

__device__ __forceinline__ float normal_variance(const DenoiseFeatures& f) { return saturate(1.f - length(f.normal)); }

// calculateFeatureWeight (denoiseFeatures.hlsl:162-176)
__device__ __forceinline__ float feature_weight(const DenoiseFeatures& a, const DenoiseFeatures& b, const DenoiseWeights& w) {
  vec3 nd = a.normal - b.normal;
  float wN = variance_weight(dot(nd, nd), w.normalSigma, normal_variance(a) + normal_variance(b), w.varianceScale);
  vec3 ad = a.albedo - b.albedo;
  float wA = variance_weight(dot(ad, ad), w.materialSigma, a.varAlbedo + b.varAlbedo, w.varianceScale);
  float m0 = a.roughness - b.roughness, m1 = a.metalness - b.metalness, m2 = a.transparency - b.transparency,
        m3 = a.anisotropy - b.anisotropy, di = a.ior - b.ior;
  float wM = variance_weight(m0 * m0 + m1 * m1 + m2 * m2 + m3 * m3 + di * di, w.materialSigma, a.varMaterial + b.varMaterial,
                             w.varianceScale);
  return wN * wA * wM;
}

// calculateRelativeDepthWeight (prefilterFeatures.comp:40-48)
__device__ __forceinline__ float relative_depth_weight(const DenoiseFeatures& c, const DenoiseFeatures& s, const DenoiseWeights& w) {
  if (c.depth <= 0.f || s.depth <= 0.f) return 1.f;
  float d = c.depth - s.depth;
  return variance_weight(d * d, c.depth * w.depthSigmaScale, c.varDepth + s.varDepth, w.varianceScale);
}

__constant__ float kGauss3[3] = {3.f / 8.f, 1.f / 4.f, 1.f / 16.f};  // precalculatedGaussian3tap

// prefilterFeatures main (prefilterFeatures.comp:50-162): moments -> means and variances, filtered
// with the variance-normalized weights (variances only between pixels of similar means)
__global__ __launch_bounds__(64) void k_prefilter_features(PostParams pp, const float4* acc, uint4* out0, uint4* out1) {
  uint32_t x = blockIdx.x * 8 + threadIdx.x % 8, y = blockIdx.y * 8 + threadIdx.x / 8;
  if (x >= pp.width || y >= pp.height) return;
  size_t npix = (size_t)pp.width * pp.height;
  uint32_t p = y * pp.width + x;
  DenoiseFeatures center = load_moments(acc, npix, p);
  if (!pp.prefilter) {
    write_features(out0, out1, p, center);
    return;
  }
  const float kGauss3x3[3] = {1.f / 4.f, 1.f / 8.f, 1.f / 16.f};
  DenoiseWeights relaxed = {pp.normalSigma, pp.materialSigma, pp.depthSigmaScale, pp.featureVarianceScale};
  DenoiseWeights strict = {pp.normalSigma, pp.materialSigma, pp.depthSigmaScale, 0.f};
  vec3 sumAlbedo(0.f), sumNormal(0.f);
  float sumMat[4] = {0.f, 0.f, 0.f, 0.f};
  float sumIor = 0.f, sumWeight = 0.f, sumDepth = 0.f, sumDepthWeight = 0.f;
  float sumVarAlbedo = 0.f, sumVarDepth = 0.f, sumVarMaterial = 0.f, sumVarWeight = 0.f;
  for (int dy = -2; dy <= 2; ++dy) {
    for (int dx = -2; dx <= 2; ++dx) {
      int sx = (int)x + dx, sy = (int)y + dy;
      if (sx < 0 || sx > (int)pp.width - 1 || sy < 0 || sy > (int)pp.height - 1) continue;
      DenoiseFeatures s = load_moments(acc, npix, (uint32_t)sy * pp.width + (uint32_t)sx);
      float kw = kGauss3[dx < 0 ? -dx : dx] * kGauss3[dy < 0 ? -dy : dy];
      float w = kw * feature_weight(center, s, relaxed) * relative_depth_weight(center, s, relaxed);
      sumAlbedo = sumAlbedo + s.albedo * w;
      sumNormal = sumNormal + s.normal * w;
      sumMat[0] += s.roughness * w; sumMat[1] += s.metalness * w; sumMat[2] += s.transparency * w; sumMat[3] += s.anisotropy * w;
      sumIor += s.ior * w;
      sumWeight += w;
      if (s.depth > 0.f) {
        sumDepth += s.depth * w;
        sumDepthWeight += w;
      }
      if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) {
        int k = (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy);
        float wv = kGauss3x3[k] * feature_weight(center, s, strict) * relative_depth_weight(center, s, strict);
        sumVarAlbedo += s.varAlbedo * wv;
        sumVarMaterial += s.varMaterial * wv;
        sumVarDepth += (s.depth > 0.f ? s.varDepth : center.varDepth) * wv;
        sumVarWeight += wv;
      }
    }
  }
  DenoiseFeatures r = center;
  float inv = 1.f / fmax_(sumWeight, 1e-6f);
  r.albedo = sumAlbedo * inv;
  r.normal = sumNormal * inv;
  r.roughness = sumMat[0] * inv; r.metalness = sumMat[1] * inv; r.transparency = sumMat[2] * inv; r.anisotropy = sumMat[3] * inv;
  r.ior = sumIor * inv;
  if (center.depth > 0.f && sumDepthWeight > 0.f) r.depth = sumDepth / sumDepthWeight;
  float invVar = 1.f / fmax_(sumVarWeight, 1e-6f);
  r.varAlbedo = sumVarAlbedo * invVar;
  r.varMaterial = sumVarMaterial * invVar;
  r.varDepth = sumVarDepth * invVar;
  write_features(out0, out1, p, r);
}

// reconstructWorldPosition (denoise.comp:104-115): planar view depth along the pinhole ray
__device__ __forceinline__ vec3 reconstruct_world(const PostParams& pp, uint32_t x, uint32_t y, float depth) {
  float in[4] = {((float)x + 0.5f) / (float)pp.width, ((float)y + 0.5f) / (float)pp.height, 0.f, 1.f}, o[4];
  mul4(pp.uvToView, in, o);
  vec3 near(o[0] / o[3], o[1] / o[3], o[2] / o[3]);
  vec3 posVS = near * (depth / fmax_(fabsf(near.z), 1e-8f));
  return xform_point(pp.viewToWorld, posVS);
}

// calculatePlaneDistanceWeight (denoise.comp:117-135)
__device__ __forceinline__ float plane_distance_weight(vec3 cPos, const DenoiseFeatures& c, vec3 sPos, const DenoiseFeatures& s,
                                                       const DenoiseWeights& w) {
  if (s.depth <= 0.f) return 1.f;
  float dist;
  float nl = length(c.normal);
  if (nl > 1e-3f) dist = dot(c.normal / nl, sPos - cPos);
  else dist = c.depth - s.depth;
  return variance_weight(dist * dist, c.depth * w.depthSigmaScale, c.varDepth + s.varDepth, w.varianceScale);
}

// XYZ -> linear sRGB (colorSpaces.hlsl c_srgbXYZToRGB) and gamutClamp (denoise.comp:272-282)
__device__ __forceinline__ float4 xyz_to_srgb_clamped(float4 c) {
  float r = 3.2406f * c.x + (-1.5372f) * c.y + (-0.4986f) * c.z;
  float g = (-0.9689f) * c.x + 1.8758f * c.y + 0.0415f * c.z;
  float b = 0.0557f * c.x + (-0.2040f) * c.y + 1.0570f * c.z;
  float lum = 0.2126f * r + 0.7152f * g + 0.0722f * b;
  float m = fmin_(fmin_(r, g), b);
  if (m < 0.f) {
    float t = safe_div(lum, lum - m);
    r = lum + t * (r - lum); g = lum + t * (g - lum); b = lum + t * (b - lum);
  }
  return make_float4(r, g, b, c.w);
}

// One pass of denoiseAtrous (denoise.comp:142-270), stride 2^pass; `on` false copies the colour
// (DENOISE_MODE_OFF). The last pass converts to linear sRGB.
__global__ __launch_bounds__(64) void k_atrous(PostParams pp, uint32_t pass, int on, int convert, const float4* src,
                                               const float* srcVar, const uint4* f0, const uint4* f1, float4* dst, float* dstVar) {
  uint32_t x = blockIdx.x * 8 + threadIdx.x % 8, y = blockIdx.y * 8 + threadIdx.x / 8;
  if (x >= pp.width || y >= pp.height) return;
  uint32_t p = y * pp.width + x;
  float4 col = src[p];
  if (on) {
    const float kGauss3x3[3][3] = {{1.f / 16, 1.f / 8, 1.f / 16}, {1.f / 8, 1.f / 4, 1.f / 8}, {1.f / 16, 1.f / 8, 1.f / 16}};
    const float EPS = 1e-6f;
    float fvs = 0.f, fvw = 0.f;
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int sx = (int)x + dx, sy = (int)y + dy;
        if (sx < 0 || sx > (int)pp.width - 1 || sy < 0 || sy > (int)pp.height - 1) continue;
        float w = kGauss3x3[dx + 1][dy + 1];
        fvs += srcVar[(uint32_t)sy * pp.width + (uint32_t)sx] * w;
        fvw += w;
      }
    float centerStd = sqrtf(fmax_(fvs / fmax_(fvw, EPS), EPS));
    DenoiseFeatures cf = read_features(f0, f1, p);
    float outVar = 0.f;
    if (cf.depth > 0.f) {
      vec3 cPos = reconstruct_world(pp, x, y, cf.depth);
      DenoiseWeights dw = {pp.normalSigma, pp.materialSigma, pp.depthSigmaScale, pp.featureVarianceScale};
      float sr = 0.f, sg = 0.f, sb = 0.f, sa = 0.f, sumVar = 0.f, sumW = 0.f;
      int step = 1 << pass;
      for (int dy = -2; dy <= 2; ++dy) {
        for (int dx = -2; dx <= 2; ++dx) {
          int sx = (int)x + dx * step, sy = (int)y + dy * step;
          if (sx < 0 || sx > (int)pp.width - 1 || sy < 0 || sy > (int)pp.height - 1) continue;
          uint32_t q = (uint32_t)sy * pp.width + (uint32_t)sx;
          float kw = kGauss3[dx < 0 ? -dx : dx] * kGauss3[dy < 0 ? -dy : dy];
          DenoiseFeatures sf = read_features(f0, f1, q);
          float4 sc = src[q];
          float sv = srcVar[q];
          vec3 sPos = reconstruct_world(pp, (uint32_t)sx, (uint32_t)sy, sf.depth);
          float wDepth = plane_distance_weight(cPos, cf, sPos, sf, dw);
          float wMat = feature_weight(cf, sf, dw);
          float wColor = expf(-fabsf(sc.y - col.y) / (pp.colorSigma * centerStd + EPS));
          float w = kw * wDepth * wMat * wColor;
          sr += sc.x * w; sg += sc.y * w; sb += sc.z * w; sa += sc.w * w;
          sumVar += w * w * sv;
          sumW += w;
        }
      }
      float inv = 1.f / fmax_(sumW, EPS);
      col = make_float4(sr * inv, sg * inv, sb * inv, sa * inv);
      outVar = sumVar * inv * inv;
    }
    dstVar[p] = outVar;
  }
  if (convert) col = xyz_to_srgb_clamped(col);
  dst[p] = col;
}

// ACESFitted (tonemapUtils.hlsl:1-36)
__device__ __forceinline__ vec3 aces_fitted(vec3 c) {
  vec3 v(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z, 0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z,
         0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z);
  vec3 a(v.x * (v.x + 0.0245786f) - 0.000090537f, v.y * (v.y + 0.0245786f) - 0.000090537f, v.z * (v.z + 0.0245786f) - 0.000090537f);
  vec3 b(v.x * (0.983729f * v.x + 0.4329510f) + 0.238081f, v.y * (0.983729f * v.y + 0.4329510f) + 0.238081f,
         v.z * (0.983729f * v.z + 0.4329510f) + 0.238081f);
  v = vec3(a.x / b.x, a.y / b.y, a.z / b.z);
  vec3 o(1.60475f * v.x + (-0.53108f) * v.y + (-0.07367f) * v.z, (-0.10208f) * v.x + 1.10813f * v.y + (-0.00605f) * v.z,
         (-0.00327f) * v.x + (-0.07276f) * v.y + 1.07602f * v.z);
  return vec3(saturate(o.x), saturate(o.y), saturate(o.z));
}

// linear -> sRGB transfer function and 8-bit quantization, as an RGBA8_SRGB target stores a value
__device__ __forceinline__ uint32_t srgb8(float c) {
  c = saturate(c);  // NaN -> 0
  float e = c <= 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.f / 2.4f) - 0.055f;
  return (uint32_t)floorf(e * 255.f + 0.5f);
}

// tonemap.frag: exposure always, the ACES operator when enabled; then the sRGB target's encoding
__global__ void k_tonemap(uint32_t npix, const float4* src, float exposure, int op, uchar4* out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  float4 c = src[p];
  vec3 v(c.x * exposure, c.y * exposure, c.z * exposure);
  float a = c.w * exposure;
  if (op) v = aces_fitted(v);
  uint32_t alpha = (uint32_t)floorf(saturate(a) * 255.f + 0.5f);  // alpha is stored linearly
  out[p] = make_uchar4((unsigned char)srgb8(v.x), (unsigned char)srgb8(v.y), (unsigned char)srgb8(v.z), (unsigned char)alpha);
}

}  // namespace ypt
