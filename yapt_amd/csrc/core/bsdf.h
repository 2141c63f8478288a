// Microfacet and auxiliary BSDF building blocks (tangent space, y up).
//
// Restates shaders/common/commonMath.hlsl:12-118, fresnel.hlsl:8-78, ggx.hlsl:11-121,
// bsdfSample.hlsl:7-437, miscBrdf.hlsl:5-203 and multiScatter.hlsl:12-238. The energy-compensation
// lookups read the directional-albedo tables of MultipleScatterLUTUtility.h (generated on the host,
// see msluts.h). `reciprocal` selects the RECIPROCAL_MULTISCATTER variant, which the reference
// compiles only into its bidirectional integrator (BDPTCommon.hlsl:6).
#pragma once
#include "common.h"
#include "ymath.h"
#include "spectral.h"
#include "msluts.h"

namespace ypt {

// ---------------------------------------------------------------- commonMath.hlsl
YPT_HD float j_reflection(vec3 wi, vec3 wm) { return safe_div(1.f, 4.f * saturate(dot(wi, wm))); }
YPT_HD float j_refraction(float eta, vec3 wo, vec3 wm, vec3 wi) {
  float dotMO = dot(wo, wm);
  float dotMI = dot(wi, wm);
  float denom = sqr(dotMO / eta + dotMI);
  return safe_div(fabsf(dotMI), denom);
}
YPT_HD bool same_hemisphere(vec3 ref, vec3 d) { return ref.y * d.y >= 0.f; }
YPT_HD vec3 force_same_hemisphere(vec3 ref, vec3 d) { return same_hemisphere(ref, d) ? d : -d; }

YPT_HD void sample_concentric_disk(float ux, float uy, float& ox, float& oy) {
  float sx = 2.f * ux - 1.f, sy = 2.f * uy - 1.f;
  if (sx == 0.f && sy == 0.f) { ox = 0.f; oy = 0.f; return; }
  float theta, r;
  if (fabsf(sx) > fabsf(sy)) { r = sx; theta = PI_OVER_4 * (sy / sx); }
  else { r = sy; theta = PI_OVER_2 - PI_OVER_4 * (sx / sy); }
  float s, c;
  ysincos(theta, &s, &c);
  ox = r * c; oy = r * s;
}

YPT_HD void construct_vector_base(vec3 v, vec3& v2, vec3& v3) {
  if (fabsf(v.x) > fabsf(v.y)) v2 = vec3(-v.z, 0.f, v.x) / sqrtf(v.x * v.x + v.z * v.z);
  else v2 = vec3(0.f, v.z, -v.y) / sqrtf(v.y * v.y + v.z * v.z);
  v3 = cross(v, v2);
}
YPT_HD vec3 make_orthogonal(vec3 to, vec3 v) { return normalize(v - to * dot(to, v)); }
// rows (t, n, b): mul(M, v) maps object -> tangent, mul(v, M) maps tangent -> object
YPT_HD mat3 basis_transform(vec3 n, vec3 t) {
  t = make_orthogonal(n, t);
  vec3 b = cross(t, n);
  return mat3(t, n, b);
}

// ---------------------------------------------------------------- fresnel.hlsl
YPT_HD float conductor_refractive_index(float r, float g) {
  float sqrtR = sqrtf(r);
  return g * ((1.f - r) / (1.f + r)) + (1.f - g) * ((1.f + sqrtR) / (1.f - sqrtR));
}
YPT_HD float conductor_extinction_sqr(float r, float n) {
  float nr = (n + 1.f) * (n + 1.f) * r - (n - 1.0f) * (n - 1.0f);
  return nr / (1.0f - r);
}
YPT_HD void conductor_n_k2(float r, float g, float& n, float& k2) {
  float rc = clampf(r, 0.0f, 0.99f);
  n = conductor_refractive_index(rc, g);
  k2 = conductor_extinction_sqr(rc, n);
}
// fresnel.hlsl:30-50, per spectral lane. The t3 term is kept exactly as the HLSL writes it.
YPT_HD float fresnel_conductor(float etaReal, float etaImg, float cosTheta) {
  float cosTheta2 = cosTheta * cosTheta;
  float sinTheta2 = 1.f - cosTheta2;
  float etaReal2 = etaReal * etaReal;
  float etaImg2 = etaImg * etaImg;
  float t0 = etaReal2 - etaImg2 - sinTheta2;
  float a2plusb2 = sqrtf(etaReal2 * etaImg2 * 4.f + t0 * t0);
  float t1 = a2plusb2 + cosTheta2;
  float a = sqrtf((a2plusb2 + t0) * 0.5f);
  float t2 = a * cosTheta2 * 2.f;
  float rs = (t1 - t2) / (t1 + t2);
  float t3 = a2plusb2 + sinTheta2 * sinTheta2 * cosTheta2;
  float t4 = t2 * sinTheta2;
  float rp = rs * (t3 - t4) / (t3 + t4);
  return (rp + rs) * 0.5f;
}
YPT_HD Spectrum fresnel_conductor(const Spectrum& er, const Spectrum& ek, float c) {
  return Spectrum(fresnel_conductor(er.s[0], ek.s[0], c), fresnel_conductor(er.s[1], ek.s[1], c),
                  fresnel_conductor(er.s[2], ek.s[2], c), fresnel_conductor(er.s[3], ek.s[3], c));
}
YPT_HD float fresnel_dielectric(float eta, float cosTheta) {  // fresnelDielectricDielectric2
  float c = saturate(cosTheta);
  float temp = eta * eta + c * c - 1.f;
  if (temp < 0.f) return 1.f;
  float g = sqrtf(temp);
  return 0.5f * sqr((g - c) / fmax_(g + c, 0.00001f)) * (1.f + sqr(((g + c) * c - 1.f) / ((g - c) * c + 1.f)));
}

// ---------------------------------------------------------------- ggx.hlsl
YPT_HD float smith_lambda_ggx(vec3 d, float ax, float ay) {
  float x2 = d.x * d.x;
  float ax2 = ax * ax;
  float y2 = d.z * d.z;
  float ay2 = ay * ay;
  float s = 1.0f + (x2 * ax2 + y2 * ay2) / (d.y * d.y);
  return (sqrtf(s) - 1.f) * 0.5f;
}
YPT_HD float g1_ggx(vec3 wo, vec3 wm, float ax, float ay) {
  float lambda = smith_lambda_ggx(wo, ax, ay);
  return (dot(wo, wm) >= 0.f ? 1.f : 0.f) / (1.0f + lambda);
}
YPT_HD float g2_ggx(vec3 wo, vec3 wi, vec3 wm, float ax, float ay) {
  float lambdaI = smith_lambda_ggx(wi, ax, ay);
  float lambdaO = smith_lambda_ggx(wo, ax, ay);
  float denom = 1.f + lambdaI + lambdaO;
  return 1.f / denom;
}
YPT_HD float d_ggx(vec3 wm, float ax, float ay) {
  float denom = (sqr(wm.x / ax) + sqr(wm.z / ay) + sqr(wm.y));
  denom *= denom;
  denom *= ax * ay * PI;
  return safe_div(1.f, denom);
}
// Heitz 2018 visible-normal sampling, y up (ggx.hlsl:60-84)
YPT_HD vec3 sample_vndf_ggx(vec3 wo, float ax, float ay, float u1, float u2) {
  vec3 Vh = normalize(vec3(ax * wo.x, ay * wo.z, wo.y));
  float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
  vec3 T1 = lensq > 0.f ? vec3(-Vh.y, Vh.x, 0.f) * (1.f / sqrtf(lensq)) : vec3(1.f, 0.f, 0.f);
  vec3 T2 = cross(Vh, T1);
  float r = sqrtf(u1);
  float phi = 2.0f * PI * u2;
  float sp, cp;
  ysincos(phi, &sp, &cp);
  float t1 = r * cp;
  float t2 = r * sp;
  float s = 0.5f * (1.0f + Vh.z);
  t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
  vec3 Nh = t1 * T1 + t2 * T2 + sqrtf(fmax_(0.0f, 1.0f - t1 * t1 - t2 * t2)) * Vh;
  return normalize(vec3(ax * Nh.x, fmax_(0.0f, Nh.z), ay * Nh.y));
}
YPT_HD float dv_ggx(vec3 wo, vec3 wm, float ax, float ay) {
  float D = d_ggx(wm, ax, ay);
  float G1 = g1_ggx(wo, wm, ax, ay);
  return safe_div(G1 * fabsf(dot(wo, wm)) * D, fabsf(wo.y));
}
YPT_HD vec3 sample_wm_ggx(vec3 wo, float ax, float ay, float u1, float u2) {
  bool flipWo = wo.y < 0.f;
  wo.y = fabsf(wo.y);
  vec3 wh = sample_vndf_ggx(wo, ax, ay, u1, u2);
  if (flipWo) { wo.y = -wo.y; wh.y = -wh.y; }
  return force_same_hemisphere(wo, wh);
}
YPT_HD float pdf_wm_ggx(vec3 wo, vec3 wm, float ax, float ay) { return dv_ggx(wo, wm, ax, ay); }

// ---------------------------------------------------------------- bsdfSample.hlsl
YPT_HD bool is_delta_ggx(float roughness) { return roughness < 0.002f; }
YPT_HD bool is_delta_ggx2(float ax, float ay) { return is_delta_ggx(fmax_(ax, ay)); }
YPT_HD bool is_delta_sheen(float roughness) { return roughness < 0.01f; }

YPT_HD vec3 sample_ggx_reflection_conductor(float ax, float ay, vec3 wo, float s0, float s1) {
  vec3 wm = is_delta_ggx2(ax, ay) ? vec3(0.f, 1.f, 0.f) : sample_wm_ggx(wo, ax, ay, s0, s1);
  vec3 wi = reflect(-wo, wm);
  if (wi.y < 0.f) return vec3(0.f);
  return wi;
}
YPT_HD float pdf_ggx_reflection_conductor(vec3 wo, vec3 wi, float ax, float ay) {
  vec3 wm = normalize(wo + wi);
  if (dot(wm, wo) < 0.f) return 0.f;
  if (is_delta_ggx2(ax, ay)) return 1.f;
  return pdf_wm_ggx(wo, wm, ax, ay) * j_reflection(wo, wm);
}
YPT_HD Spectrum eval_ggx_reflection_conductor(const Spectrum& etaR, const Spectrum& etaK, float ax, float ay, vec3 wo, vec3 wi) {
  vec3 wm = normalize(wo + wi);
  if (!same_hemisphere(wo, wi)) return Spectrum(0.f);
  Spectrum F = fresnel_conductor(etaR, etaK, dot(wo, wm));
  if (is_delta_ggx2(ax, ay)) return F / fabsf(wi.y);
  float G2 = g2_ggx(wo, wi, wm, ax, ay);
  float D = d_ggx(wm, ax, ay);
  return F * G2 * D / fmax_(4.f * wo.y * wi.y, 0.00001f);
}

YPT_HD vec3 sample_ggx_reflection_dielectric(float ax, float ay, vec3 wo, float s0, float s1) {
  vec3 wm = is_delta_ggx2(ax, ay) ? vec3(0.f, 1.f, 0.f) : sample_wm_ggx(wo, ax, ay, s0, s1);
  vec3 wi = reflect(-wo, wm);
  if (!same_hemisphere(wo, wi)) return vec3(0.f);
  return wi;
}
YPT_HD float pdf_ggx_reflection_dielectric(vec3 wo, vec3 wi, float ax, float ay) {
  if (!same_hemisphere(wo, wi)) return 0.f;
  if (is_delta_ggx2(ax, ay)) return 1.f;
  vec3 wm = normalize(wo + wi);
  return pdf_wm_ggx(wo, wm, ax, ay) * j_reflection(wo, wm);
}
YPT_HD float eval_ggx_reflection_dielectric(float etaR, float ax, float ay, vec3 wo, vec3 wi) {
  vec3 wm = normalize(wo + wi);
  if (!same_hemisphere(wo, wi)) return 0.f;
  if (dot(wi, wm) > 0.f && dot(wo, wm) > 0.f) {
    float F = fresnel_dielectric(etaR, dot(wo, wm));
    if (is_delta_ggx2(ax, ay)) return F / fabsf(wi.y);
    float D = d_ggx(wm, ax, ay);
    float G2 = g2_ggx(wo, wi, wm, ax, ay);
    return F * G2 * D / fmax_(4.f * wo.y * wi.y, 0.00001f);
  }
  return 0.f;
}

YPT_HD vec3 wm_translucent(vec3 wo, vec3 wi, float etaR) {
  vec3 wm;
  if (same_hemisphere(wo, wi)) {
    wm = normalize(wo + wi);
  } else {
    if (etaR == 1.f) etaR = 1.001f;
    wm = normalize(wo + wi * etaR);
    wm = force_same_hemisphere(wo, wm);
  }
  return wm;
}
YPT_HD float pdf_ggx_transmitted(float etaR, vec3 wo, vec3 wi, float ax, float ay) {
  if (etaR == 1.f) etaR = 1.001f;
  vec3 wm = wm_translucent(wo, wi, etaR);
  if (same_hemisphere(wo, wi)) return 0.f;
  if (is_delta_ggx2(ax, ay)) return 1.f;
  float pdf = pdf_wm_ggx(wo, wm, ax, ay);
  if (dot(wo, wm) <= 0.f || dot(wi, wm) >= 0.f) return 0.f;
  pdf *= j_refraction(etaR, wo, wm, wi);
  return pdf;
}
YPT_HD float eval_ggx_transmitted(float etaR, float ax, float ay, vec3 wo, vec3 wi) {
  if (etaR == 1.f) etaR = 1.001f;
  vec3 wm = wm_translucent(wo, wi, etaR);
  if (same_hemisphere(wo, wi)) return 0.f;
  if (dot(wo, wm) <= 0.f || dot(wi, wm) >= 0.f) return 0.f;
  float VdotH = fabsf(dot(wo, wm));
  float F = fresnel_dielectric(etaR, VdotH);
  if (is_delta_ggx2(ax, ay)) return (1.f - F) / fabsf(wi.y);
  float G2 = g2_ggx(wo, wi, wm, ax, ay);
  float D = d_ggx(wm, ax, ay);
  float denom = fabsf(wo.y * wi.y);
  float weight = (1.f - F) * G2 * D * VdotH * j_refraction(etaR, wo, wm, wi);
  return safe_div(weight, denom);
}

YPT_HD void refl_refr_probabilities(float etaR, vec3 wo, vec3 wm, bool allowRefl, bool allowTrans, float& pr, float& pt) {
  float VdotH = saturate(dot(wo, wm));
  float reflProb = fresnel_dielectric(etaR, VdotH);
  float refrProb = 1.f - reflProb;
  if (!allowRefl) reflProb = 0.f;
  if (!allowTrans) refrProb = 0.f;
  if (reflProb == 0.f && refrProb == 0.f) { pr = 0.f; pt = 0.f; return; }
  float sum = reflProb + refrProb;
  pr = safe_div(reflProb, sum);
  pt = safe_div(refrProb, sum);
}
YPT_HD vec3 sample_dielectric(float etaR, float ax, float ay, vec3 wo, float s0, float s1, float sc, bool allowRefl, bool allowTrans) {
  vec3 wm;
  if (is_delta_ggx2(ax, ay)) wm = wo.y < 0.f ? vec3(0.f, -1.f, 0.f) : vec3(0.f, 1.f, 0.f);
  else wm = sample_wm_ggx(wo, ax, ay, s0, s1);
  if (dot(wo, wm) < 0.f) return vec3(0.f);
  if (etaR == 1.f) etaR = 1.001f;
  float pr, pt;
  refl_refr_probabilities(etaR, wo, wm, allowRefl, allowTrans, pr, pt);
  if (pr == 0.f && pt == 0.f) return vec3(0.f);
  if (sc < pr || pt == 0.f) {
    vec3 wi = reflect(-wo, wm);
    if (!same_hemisphere(wo, wi)) return vec3(0.f);
    return wi;
  }
  float invEta = 1.f / etaR;
  vec3 wi = refract(-wo, wm, invEta);
  if (same_hemisphere(wo, wi)) return vec3(0.f);
  return wi;
}
YPT_HD float pdf_dielectric(float etaR, vec3 wo, vec3 wi, float ax, float ay, bool allowRefl, bool allowTrans) {
  vec3 wm = wm_translucent(wo, wi, etaR);
  float pr, pt;
  refl_refr_probabilities(etaR, wo, wm, allowRefl, allowTrans, pr, pt);
  if (same_hemisphere(wo, wi)) return pdf_ggx_reflection_dielectric(wo, wi, ax, ay) * pr;
  return pdf_ggx_transmitted(etaR, wo, wi, ax, ay) * pt;
}

// ---------------------------------------------------------------- miscBrdf.hlsl
YPT_HD vec3 sample_cosine_hemisphere(float r0, float r1) {
  float y = sqrtf(1.f - r0);
  float r = sqrtf(r0);
  float phi = 2.0f * PI * r1;
  float s, c;
  ysincos(phi, &s, &c);
  return vec3(r * c, y, r * s);
}
YPT_HD vec3 sample_cosine_power_hemisphere(float r0, float r1, float exponent) {
  float y = ypow(1.f - r0, 1.f / (exponent + 2.f));
  float r = sqrtf(fmax_(0.f, 1.f - sqr(y)));
  float phi = 2.0f * PI * r1;
  float s, c;
  ysincos(phi, &s, &c);
  return vec3(r * c, y, r * s);
}
YPT_HD vec3 sample_cosine_power_sphere(float r0, float r1, float exponent) {
  bool flip = r0 < 0.5f;
  r0 = flip ? r0 * 2.f : (r0 - 0.5f) * 2.f;
  vec3 d = sample_cosine_power_hemisphere(r0, r1, exponent);
  d.y = flip ? -d.y : d.y;
  return d;
}
YPT_HD float pdf_cosine_hemisphere(float dotIN) { return saturate(dotIN) * INVPI; }
YPT_HD float pdf_cosine_power_hemisphere(float dotIN, float exponent) {
  return (exponent + 2.f) * INVPI * 0.5f * ypow(saturate(dotIN), exponent + 1.f);
}
YPT_HD float pdf_cosine_power_sphere(float dotIN, float exponent) {
  return (exponent + 2.f) * INVPI * 0.25f * ypow(fabsf(dotIN), exponent + 1.f);
}
// Le(w) = emissive * (n+2)/2 * |cos|^n   (miscBrdf.hlsl:124-130)
YPT_HD float emission_profile(float dotIN, float exponent) {
  float profile = exponent > 0.f ? ypow(fabsf(dotIN), exponent) : 1.f;
  return (exponent + 2.f) * 0.5f * profile;
}
YPT_HD float gl_sheen(float x, float a, float b, float c, float d, float e) {
  return (a / (1.f + b * ypow(x, c))) + d * x + e;
}
YPT_HD float glambda_sheen(float roughness, float cosTheta) {  // Imageworks sheen fit
  const float a0 = 25.3245f, b0 = 3.32435f, c0 = 0.16801f, d0 = -1.27393f, e0 = -4.85967f;
  const float a1 = 21.5473f, b1 = 3.82987f, c1 = 0.19823f, d1 = -1.97760f, e1 = -4.32054f;
  float t = 1.f - roughness;
  t *= t;
  float a = lerpf(a1, a0, t), b = lerpf(b1, b0, t), c = lerpf(c1, c0, t), d = lerpf(d1, d0, t), e = lerpf(e1, e0, t);
  float exponent;
  if (cosTheta < 0.5f) exponent = gl_sheen(cosTheta, a, b, c, d, e);
  else exponent = 2.f * gl_sheen(0.5f, a, b, c, d, e) - gl_sheen(1.f - cosTheta, a, b, c, d, e);
  return yexp(exponent);
}
YPT_HD float d_sheen(vec3 wm, float r) {
  float rInv = 1.f / r;
  float sinTheta2 = fmax_(1.f - wm.y * wm.y, 0.0001f);
  float v = (2.f + rInv) * ypow(sinTheta2, rInv * 0.5f);
  return v / (2.f * PI);
}
YPT_HD float g_sheen(vec3 wo, vec3 wi, float r) {
  float lambdaI = glambda_sheen(r, wi.y);
  float lambdaO = glambda_sheen(r, wo.y);
  return 1.f / (1.f + lambdaI + lambdaO);
}
YPT_HD vec3 sample_sheen(float r, vec3 wo, float s0, float s1) {
  vec3 wi = is_delta_sheen(r) ? reflect(-wo, vec3(0.f, 1.f, 0.f)) : sample_cosine_hemisphere(s0, s1);
  if (wi.y < 0.f) return vec3(0.f);
  return wi;
}
YPT_HD float pdf_sheen(vec3 wo, vec3 wi, float r) {
  if (wi.y < 0.0f) return 0.f;
  return is_delta_sheen(r) ? 1.f : pdf_cosine_hemisphere(wi.y);
}
YPT_HD Spectrum eval_sheen(const Spectrum& sheenColor, float r, vec3 wo, vec3 wi) {
  if (wi.y < 0.0f) return Spectrum(0.f);
  vec3 wm = normalize(wo + wi);
  if (dot(wi, wm) > 0.f && dot(wo, wm) > 0.f) {
    if (is_delta_sheen(r)) {
      float d = fabsf(wi.y);
      return d == 0.f ? Spectrum(0.f) : sheenColor / d;
    }
    float D = d_sheen(wm, r);
    float G = g_sheen(wo, wi, r);
    float DGDenom = fabsf(safe_div(D * G, 4.f * wo.y * wi.y));
    return sheenColor * DGDenom;
  }
  return Spectrum(0.f);
}
YPT_HD vec3 sample_diffuse_lambertian(vec3 wo, float s0, float s1) {
  vec3 wi = sample_cosine_hemisphere(s0, s1);
  if (wi.y < 0.f) return vec3(0.f);
  return wi;
}
YPT_HD float pdf_diffuse_lambertian(vec3 wi) {
  if (wi.y < 0.0f) return 0.f;
  return pdf_cosine_hemisphere(wi.y);
}

// ---------------------------------------------------------------- multiScatter.hlsl
YPT_HD float directional_albedo_translucent(const MSLuts& L, float etaR, float dotDir, float lr) {
  return etaR < 1.f ? L.tr_dir_lighter(etaR, fabsf(dotDir), lr) : L.tr_dir_denser(etaR, fabsf(dotDir), lr);
}
YPT_HD float average_albedo_translucent(const MSLuts& L, float etaR, float lr) {
  return etaR < 1.f ? L.tr_avg_lighter(etaR, lr) : L.tr_avg_denser(etaR, lr);
}
YPT_HD float avg_fresnel(float etaR) {
  if (etaR < 1.f) return 0.997118f + 0.1014f * etaR - 0.965241f * etaR * etaR - 0.130607f * etaR * etaR * etaR;
  return (etaR - 1.f) / (4.08567f + 1.00071f * etaR);
}
YPT_HD float ms_transmission_ratio(const MSLuts& L, float etaR, float lr) {
  float etaOut = fmax_(etaR, 1.f / etaR);
  float tOut = 1.f - avg_fresnel(etaOut);
  if (etaR >= 1.f) return tOut;
  float eAvgOut = average_albedo_translucent(L, etaOut, lr);
  float eAvgIn = average_albedo_translucent(L, etaR, lr);
  return saturate(safe_div(tOut * (1.f - eAvgOut), etaOut * etaOut * (1.f - eAvgIn)));
}
YPT_HD float fms_dielectric(const MSLuts& L, bool reciprocal, float etaR, float lr, float cosTheta) {
  if (reciprocal) {
    float fAvg = avg_fresnel(etaR);
    float eAvg = L.ss_avg(lr);
    return fAvg * eAvg / (1.f - fAvg * (1.f - eAvg));
  }
  return fresnel_dielectric(etaR, cosTheta);
}
// getEnergyCompensation for a scalar single-scatter term
YPT_HD float energy_compensation(const MSLuts& L, bool reciprocal, float fms, float dotWo, float dotWi, float lr, float ss) {
  dotWo = fabsf(dotWo); dotWi = fabsf(dotWi);
  if (reciprocal) {
    float eo = L.ss_dir(fabsf(dotWo), lr);
    float ei = L.ss_dir(fabsf(dotWi), lr);
    float eAvg = L.ss_avg(lr);
    float ems = safe_div((1.f - eo) * (1.f - ei), (1.f - eAvg) * PI);
    return fms * ems;
  }
  float e = L.ss_dir(fabsf(dotWo), lr);
  float k = fms * (1.f - e) / e;
  return ss * k;
}
YPT_HD Spectrum energy_compensation(const MSLuts& L, bool reciprocal, const Spectrum& fms, float dotWo, float dotWi, float lr, const Spectrum& ss) {
  dotWo = fabsf(dotWo); dotWi = fabsf(dotWi);
  if (reciprocal) {
    float eo = L.ss_dir(fabsf(dotWo), lr);
    float ei = L.ss_dir(fabsf(dotWi), lr);
    float eAvg = L.ss_avg(lr);
    float ems = safe_div((1.f - eo) * (1.f - ei), (1.f - eAvg) * PI);
    return fms * ems;
  }
  float e = L.ss_dir(fabsf(dotWo), lr);
  Spectrum k = fms * (1.f - e);
  k = k / e;
  return ss * k;
}
YPT_HD float energy_remaining_after_specular(const MSLuts& L, float etaR, float dotWo, float lr, float amount) {
  return 1.f - L.ssms_dir(etaR, fabsf(dotWo), lr) * amount;
}
YPT_HD float energy_remaining_after_sheen(const MSLuts& L, float dotWo, float lr, float amount) {
  return 1.f - L.sheen_dir(fabsf(dotWo), lr) * amount;
}
YPT_HD float energy_compensation_translucent(const MSLuts& L, bool reciprocal, float etaR, float dotWo, float dotWi, float lr, float ss) {
  if (reciprocal) {
    bool transmission = dotWo * dotWi < 0.f;
    float dWo, dWi, eAvg, ratio;
    float tr = ms_transmission_ratio(L, etaR, lr);
    if (transmission) {
      float etaInv = 1.f / etaR;
      dWo = directional_albedo_translucent(L, etaR, dotWo, lr);
      dWi = directional_albedo_translucent(L, etaInv, dotWi, lr);
      eAvg = average_albedo_translucent(L, etaInv, lr);
      ratio = tr;
    } else {
      dWo = directional_albedo_translucent(L, etaR, dotWo, lr);
      dWi = directional_albedo_translucent(L, etaR, dotWi, lr);
      eAvg = average_albedo_translucent(L, etaR, lr);
      ratio = 1.f - tr;
    }
    return ratio * safe_div((1.f - dWo) * (1.f - dWi), (1.f - eAvg) * PI);
  }
  float e = fmax_(directional_albedo_translucent(L, etaR, fabsf(dotWo), lr), 0.1f);  // MIN_TRANSLUCENT_DIRECTIONAL_ALBEDO
  return (ss / e) - ss;
}

}  // namespace ypt
