// Layered principled material: surface definition, layers, sampling and evaluation.
//
// Restates shaders/materials/materialsCommon.hlsl:22-568 (surface parameters, delta test,
// regularisation, two-sided orientation, anisotropy rotation, Cauchy dispersion, thin-film
// interference), shaders/materials/materialLayers.hlsl:16-443 (coating, sheen, conductor, diffuse and
// specular layers) and shaders/raytracing/materialSample.hlsl:89-611 (layer selection, direction
// sampling, evaluateSurface). Subsurface scattering (LAYERIND_SUBSURFACE) is selected with its
// probability but yields no direction: the configurations here use subsurfaceAmount = 0.
#pragma once
#include "bsdf.h"

namespace ypt {

static constexpr uint32_t MATMASK_TWO_SIDED = 1u << 0;   // MaterialMask_TwoSided
static constexpr uint32_t MATMASK_DISPERSION = 1u << 1;  // MaterialMask_Dispersion
static constexpr uint32_t MATMASK_ALPHA_BLEND = 1u << 2;

// Material record as the kernels read it: the parameters of MaterialEntryGPU
// (materialsCommonResources.hlsl:10-43, packed with the clamps of BindlessMaterialManager.cpp:70-100)
// with every colour already converted to sigmoid-polynomial coefficients on the host.
struct MaterialGPU {
  SpectralColor specular, albedo, absorption, emissive, sheenColor, subsurfaceRadius;
  float specularAmount, clearCoatAmount, clearCoatIOR, clearCoatRoughness;
  float transparency, metalness, dielectricIOR, roughness;
  float anisotropy, anisotropyRotation, thinFilmThicknessNM, sheenRoughness;
  float sheenAmount, emissionFocus, subsurfaceAmount, alphaCutoff;
  float cauchyA, cauchyB;
  uint32_t flags;
  uint32_t emissiveNonZero;
  // linear RGB: the first-hit feature buffers, and the factors textures multiply
  float albedoRGB[3], emissiveRGB[3], sheenRGB[3];
  float blackbodyK;                     // > 0: Planck emission spectrum (an extension)
  float sellmeierB[3], sellmeierC[3];   // sellmeierB[0] > 0: Sellmeier dispersion (an extension)
  // textures (albedo, normal, occlusion-roughness-metalness, emissive): index or -1, UV scale
  int32_t tex[4];
  float texScale[8];
};

static constexpr int TEX_ALBEDO = 0, TEX_NORMAL = 1, TEX_ORM = 2, TEX_EMISSIVE = 3;

// Texture values at a surface point, sampled by the caller (scene.h sample_surface_texels), and the
// RGB -> spectrum table that textured colours go through at run time.
struct SurfaceTexels {
  uint32_t mask;   // bit k: texture k is bound and was sampled
  float v[4][4];   // RGBA per texture, linear
  const float* rgb2spec;
};

struct SurfaceDef {
  Spectrum specular, albedo, absorption, emissive, sheenColor;
  vec3 coatingNormal, baseNormal, geometryNormal, tangent;
  float transparency, metalness, dielectricIOR, roughness;
  float dispersedIOR, specularAmount, clearCoatAmount;  // dispersedIOR: at the hero wavelength
  float clearCoatIOR, clearCoatRoughness, anisotropy, anisotropyRotation;
  float thinFilmThicknessNM, sheenRoughness, sheenAmount, emissionFocus;
  float subsurfaceAmount, diffOcclusion, specOcclusion;
  uint32_t flags;
};

YPT_HD bool surface_two_sided(uint32_t flags) { return (flags & MATMASK_TWO_SIDED) != 0; }
YPT_HD bool has_dispersion(uint32_t flags) { return (flags & MATMASK_DISPERSION) != 0; }

// calculateRoughnessParams (surfaceParameters.hlsl:6-10)
YPT_HD void roughness_params(float roughness, float anisotropy, float& ax, float& ay) {
  float a2 = roughness * roughness;
  ax = clampf(a2 * (1.f + anisotropy), 0.0f, 1.f);
  ay = clampf(a2 * (1.f - anisotropy), 0.0f, 1.f);
}

YPT_HD float regularize(float linearRoughness, bool squared) {  // materialsCommon.hlsl:399-411
  float r = squared ? linearRoughness * linearRoughness : linearRoughness;
  if (r < 0.3f) {
    r = clampf(r * 2.f, 0.1f, 0.3f);
    return squared ? sqrtf(r) : r;
  }
  return linearRoughness;
}

YPT_HD vec3 orient_to_same_side(vec3 wo, vec3 n) { return dot(wo, n) < 0.f ? -n : n; }

YPT_HD float refractive_index_for_wavelength(float A, float B, float lambdaNM) {  // Cauchy, :503-511
  float um = lambdaNM * 0.001f;
  return A + (B / (um * um));
}

// Sellmeier's equation, n^2 = 1 + sum B_i l^2 / (l^2 - C_i) with l in micrometres: the dispersion the
// benchmark glass names (BK7), an extension next to the reference's Cauchy form
YPT_HD float sellmeier_ior(const float B[3], const float C[3], float lambdaNM) {
  float l2 = lambdaNM * 0.001f * (lambdaNM * 0.001f);
  float n2 = 1.f + B[0] * l2 / (l2 - C[0]) + B[1] * l2 / (l2 - C[1]) + B[2] * l2 / (l2 - C[2]);
  return sqrtf(fmax_(n2, 1e-6f));
}

// the IOR a dispersive material has at the hero wavelength
YPT_HD float material_dispersed_ior(const MaterialGPU& m, float heroNM) {
  return m.sellmeierB[0] > 0.f ? sellmeier_ior(m.sellmeierB, m.sellmeierC, heroNM)
                               : refractive_index_for_wavelength(m.cauchyA, m.cauchyB, heroNM);
}

// Planck's law relative to its value at 560 nm: (560 / l)^5 (e^(c2 / (560 nm T)) - 1) / (e^(c2 / (l T)) - 1)
YPT_HD float planck_relative(float lambdaNM, float kelvin) {
  const float c2 = 1.4387769e7f;  // h c / k in nm K
  float a = 560.f / lambdaNM;
  float num = yexp(c2 / (560.f * kelvin)) - 1.f;
  float den = yexp(c2 / (lambdaNM * kelvin)) - 1.f;
  return a * a * a * a * a * (num / den);
}
YPT_HD Spectrum blackbody_spectrum(const Lambdas& L, float kelvin, float strength) {
  Spectrum r;
  for (int i = 0; i < 4; ++i) r.s[i] = strength * planck_relative(L.l[i], kelvin);
  return r;
}

// thin film (materialsCommon.hlsl:515-568)
YPT_HD float optical_path_difference(float etaR, float cosIncident, float nFilm, float thickness) {
  float cosRefractedSqr = 1.f - (1.f - cosIncident * cosIncident) / (etaR * etaR);
  if (cosRefractedSqr < 0.f) return 0.f;
  return 2.f * nFilm * thickness * sqrtf(cosRefractedSqr);
}
YPT_HD float thin_film_interference_opd(float waveLambda, float opd, float phaseShift) {
  float m = opd / waveLambda + phaseShift;
  return ycos(2.f * m * PI);
}
YPT_HD void thin_film_interference(const Lambdas& L, float nFilm, float etaRTop, float phaseShift, float cosIncident,
                                   float thickness, float out[4]) {
  const float INCOHERENCE_THRESHOLD = 1200.f;
  const float INCOHERENCE_FADE = 200.f;
  const float INCOHERENCE_FADE_INV = 1.f / 200.f;
  if (thickness > INCOHERENCE_THRESHOLD) { out[0] = out[1] = out[2] = out[3] = 0.f; return; }
  float opd = optical_path_difference(etaRTop, cosIncident, nFilm, thickness);
  float t = fmax_(0.f, thickness - INCOHERENCE_THRESHOLD + INCOHERENCE_FADE) * INCOHERENCE_FADE_INV;
  for (int i = 0; i < 4; ++i) {
    float v = thin_film_interference_opd(L.l[i], opd, phaseShift);
    out[i] = v + t * (0.f - v);
  }
}
YPT_HD Spectrum apply_thin_film(const Lambdas& L, const Spectrum& A1, const Spectrum& A2, float nFilm, float etaRTop,
                                float phaseShift, float cosIncident, float thickness) {
  bool a1any = !A1.all_equal(0.f), a2any = !A2.all_equal(0.f);
  if (!a1any) return A2;
  if (!a2any) return A1;
  float m[4];
  thin_film_interference(L, nFilm, etaRTop, phaseShift, cosIncident, thickness, m);
  Spectrum r;
  for (int i = 0; i < 4; ++i) r.s[i] = fmax_(0.f, A1.s[i] + A2.s[i] + 2.f * sqrtf(fmax_(0.f, A1.s[i] * A2.s[i])) * m[i]);
  return r;
}

// isDeltaDistribution (materialsCommon.hlsl:65-80)
YPT_HD bool is_delta_distribution(float ax, float ay, const SurfaceDef& s) {
  if (s.metalness == 1.f || s.transparency == 1.f) {
    bool coatingIsRough = (s.clearCoatAmount > 0.f && !is_delta_ggx(s.clearCoatRoughness)) ||
                          (s.sheenAmount > 0.f && !is_delta_sheen(s.sheenRoughness));
    bool baseIsRough = !is_delta_ggx2(ax, ay);
    return !coatingIsRough && !baseIsRough;
  }
  return false;
}

// fetchSurfaceMaterialParameters + handleTwoSidedMaterialOrientation + setupSurfaceOrientation +
// (optional) regularizeMaterial + convertSurfaceDefinitionFromRGB, for untextured materials.
// hitShadersCommon.hlsl:51-91, materialsCommon.hlsl:313-352, 413-418, 440-448, 487-501.
YPT_HD SurfaceDef make_surface(const MaterialGPU& m, vec3 geometryNormal, vec3 shadingNormal, vec3 tangent,
                               vec3 wo, bool applyTwoSided, bool regularizePath, const Lambdas& L,
                               const SpectralTables& T, const SurfaceTexels* tx = nullptr) {
  SurfaceDef s;
  s.transparency = m.transparency;
  s.metalness = m.metalness;
  s.dielectricIOR = m.dielectricIOR;
  s.roughness = m.roughness;
  s.specularAmount = m.specularAmount;
  s.clearCoatAmount = m.clearCoatAmount;
  s.clearCoatIOR = m.clearCoatIOR;
  s.clearCoatRoughness = m.clearCoatRoughness;
  s.dispersedIOR = has_dispersion(m.flags) ? material_dispersed_ior(m, L.l[0]) : 0.f;
  s.anisotropy = m.anisotropy;
  s.anisotropyRotation = m.anisotropyRotation;
  s.thinFilmThicknessNM = m.thinFilmThicknessNM;
  s.flags = m.flags;
  s.sheenRoughness = fmax_(m.sheenRoughness, 0.07f);
  s.sheenAmount = m.sheenAmount;
  s.emissionFocus = m.emissionFocus;
  s.subsurfaceAmount = m.subsurfaceAmount;
  s.diffOcclusion = 1.f;
  s.specOcclusion = 1.f;
  s.coatingNormal = shadingNormal;
  s.baseNormal = shadingNormal;
  s.geometryNormal = geometryNormal;
  s.tangent = tangent;
  // modifySurfaceMaterialParametersWithTextures (materialsCommon.hlsl:354-397), on the RGB values
  SpectralColor albedoCol = m.albedo, emissiveCol = m.emissive;
  if (tx && tx->mask) {
    if (tx->mask & (1u << TEX_ALBEDO)) {
      const float* a = tx->v[TEX_ALBEDO];
      vec3 rgb(m.albedoRGB[0] * a[0], m.albedoRGB[1] * a[1], m.albedoRGB[2] * a[2]);
      rgb_coeffs_table(tx->rgb2spec, rgb, albedoCol.c);
      albedoCol.scale = 1.f;
      if (s.flags & MATMASK_ALPHA_BLEND) s.transparency = 1.f - a[3] * (1.f - s.transparency);
    }
    if (tx->mask & (1u << TEX_NORMAL)) {
      const float* t = tx->v[TEX_NORMAL];
      vec3 n = normalize(vec3(t[0] * 2.f - 1.f, t[1] * 2.f - 1.f, t[2] * 2.f - 1.f));
      mat3 tbase = basis_transform(s.baseNormal, s.tangent);
      s.baseNormal = mul(vec3(n.x, n.z, n.y), tbase);
    }
    if (tx->mask & (1u << TEX_ORM)) {
      const float* o = tx->v[TEX_ORM];
      float occ = f16_to_f32(f32_to_f16(o[0]));  // packOcclusion / unpackOcclusion: half precision
      s.diffOcclusion = occ;
      s.specOcclusion = occ;
      s.roughness *= o[1];
      s.metalness *= o[2];
    }
    if (tx->mask & (1u << TEX_EMISSIVE)) {
      const float* e = tx->v[TEX_EMISSIVE];
      vec3 rgb(m.emissiveRGB[0] * e[0], m.emissiveRGB[1] * e[1], m.emissiveRGB[2] * e[2]);
      float mx = fmax_(fmax_(rgb.x, rgb.y), rgb.z);
      float scale = 2.f * mx;
      vec3 nv = scale == 0.f ? vec3(0.f) : rgb / scale;
      rgb_coeffs_table(tx->rgb2spec, nv, emissiveCol.c);
      emissiveCol.scale = scale;
    }
  }
  if (applyTwoSided && s.transparency == 0.f && surface_two_sided(s.flags)) {
    s.baseNormal = orient_to_same_side(wo, s.baseNormal);
    s.coatingNormal = orient_to_same_side(wo, s.coatingNormal);
    s.geometryNormal = orient_to_same_side(wo, s.geometryNormal);
  }
  if (s.anisotropyRotation > 0.f) {
    float rotAngle = s.anisotropyRotation * 2.f * PI;
    float sinA, cosA;
    ysincos(rotAngle, &sinA, &cosA);
    mat3 rot(vec3(cosA, 0.f, sinA), vec3(0.f, 1.f, 0.f), vec3(-sinA, 0.f, cosA));
    s.tangent = mul(rot, s.tangent);
  }
  if (regularizePath) {
    s.roughness = regularize(s.roughness, true);
    s.sheenRoughness = regularize(s.sheenRoughness, false);
    s.clearCoatRoughness = regularize(s.clearCoatRoughness, true);
  }
  s.specular = spectrum_bounded(m.specular, L);
  s.albedo = spectrum_bounded(albedoCol, L);
  s.absorption = spectrum_unbounded_no_illum(m.absorption, L);
  if (m.blackbodyK > 0.f) {  // Planck spectrum scaled by the largest (textured) emissive component
    float e0 = m.emissiveRGB[0], e1 = m.emissiveRGB[1], e2 = m.emissiveRGB[2];
    if (tx && (tx->mask & (1u << TEX_EMISSIVE))) {
      e0 *= tx->v[TEX_EMISSIVE][0]; e1 *= tx->v[TEX_EMISSIVE][1]; e2 *= tx->v[TEX_EMISSIVE][2];
    }
    s.emissive = blackbody_spectrum(L, m.blackbodyK, fmax_(fmax_(e0, e1), e2));
  } else {
    s.emissive = spectrum_unbounded(emissiveCol, L, T);
  }
  s.sheenColor = spectrum_bounded(m.sheenColor, L);
  return s;
}

// ---------------------------------------------------------------- materialSample.hlsl
static constexpr int LAYER_COATING_GGX = 0;
static constexpr int LAYER_COATING_SHEEN = 1;
static constexpr int LAYER_SPEC_CONDUCTOR = 2;
static constexpr int LAYER_SPEC_DIELECTRIC = 3;
static constexpr int LAYER_DIFFUSE_REFL = 4;
static constexpr int LAYER_SUBSURFACE = 5;
static constexpr int LAYER_COUNT = 6;

static constexpr uint32_t TRANSMISSION_NONE = 0;
static constexpr uint32_t TRANSMISSION_ENTERED = 1;
static constexpr uint32_t TRANSMISSION_EXITED = 2;
static constexpr uint32_t TRANSMISSION_DISPERSED = 4;

static constexpr uint32_t EVAL_NONE = 1u << 0;
static constexpr uint32_t EVAL_PDF_ONLY = 1u << 1;
static constexpr uint32_t EVAL_LIGHT_PATH = 1u << 2;
static constexpr uint32_t EVAL_TREAT_AS_DELTA = 1u << 3;
static constexpr uint32_t EVAL_DONT_PROJECT_SOLID_ANGLE = 1u << 4;

struct PrecalcSurface {
  vec3 coatingNormal, baseNormal, woBase, woCoating;
  float fromIOR, toIOR, ax, ay;
  bool exiting;
  YPT_HD mat3 to_coating(vec3 tangent) const { return basis_transform(coatingNormal, tangent); }
  YPT_HD mat3 to_base(vec3 tangent) const { return basis_transform(baseNormal, tangent); }
};

struct SurfaceContext {
  SurfaceDef def;
  PrecalcSurface pre;
  float prob[LAYER_COUNT];
  Lambdas lambdas;
  bool reciprocal;  // RECIPROCAL_MULTISCATTER (bidirectional integrator only)
};

struct SurfaceEval {
  Spectrum weight;
  float pdf;
  uint32_t transmission;
};

YPT_HD bool considered_diffuse_ggx(float ax, float ay) { return fmin_(ax, ay) > 0.05f; }
YPT_HD bool considered_diffuse_sheen(float r) { return r > 0.15f; }

YPT_HD void layer_probabilities(const SurfaceDef& s, const PrecalcSurface& p, float prob[LAYER_COUNT]) {
  bool comingFromInside = p.woBase.y < 0.f;
  float fromInside = comingFromInside ? 0.f : 1.f;
  float dielAmount = 1.f - s.metalness;
  float opaqueDiel = dielAmount * (1.f - s.transparency);
  float sum = 0.f;
  prob[LAYER_COATING_GGX] = s.clearCoatAmount * fromInside; sum += prob[0];
  prob[LAYER_COATING_SHEEN] = s.sheenAmount * fromInside; sum += prob[1];
  prob[LAYER_SPEC_CONDUCTOR] = s.metalness * fromInside; sum += prob[2];
  prob[LAYER_SPEC_DIELECTRIC] = dielAmount * (s.specularAmount + s.transparency); sum += prob[3];
  prob[LAYER_DIFFUSE_REFL] = opaqueDiel * (1.f - s.subsurfaceAmount) * fromInside; sum += prob[4];
  prob[LAYER_SUBSURFACE] = opaqueDiel * s.subsurfaceAmount * fromInside; sum += prob[5];
  sum = fmax_(0.000001f, sum);
  for (int i = 0; i < LAYER_COUNT; ++i) prob[i] /= sum;
}

YPT_HD int select_layer(float r, const float prob[LAYER_COUNT]) {
  float t = fmax_(0.f, r + 0.00001f);
  int layer;
  for (layer = 0; layer < LAYER_COUNT; ++layer) {
    if (t < prob[layer] && prob[layer] > 0.f) break;
    t -= prob[layer];
  }
  return layer;
}

// getPrecalculatedSurfaceData + makeSurfaceSamplingContext (:564-611)
YPT_HD void make_surface_context(SurfaceContext& c, const SurfaceDef& s, float currentIOR, float previousIOR, vec3 woObj,
                                 bool frontFace, const Lambdas& L, bool reciprocal) {
  c.def = s;
  c.lambdas = L;
  c.reciprocal = reciprocal;
  PrecalcSurface& p = c.pre;
  roughness_params(s.roughness, s.anisotropy, p.ax, p.ay);
  p.exiting = !frontFace;
  p.fromIOR = currentIOR;
  p.toIOR = p.exiting ? previousIOR : s.dielectricIOR;
  p.baseNormal = s.baseNormal;
  p.coatingNormal = s.coatingNormal;
  vec3 woC = mul(p.to_coating(s.tangent), woObj);
  vec3 woB = mul(p.to_base(s.tangent), woObj);
  p.woCoating = normalize(woC);
  p.woBase = normalize(woB);
  layer_probabilities(s, p, c.prob);
}

// --- layers (materialLayers.hlsl) ---
YPT_HD float coat_eval(const MSLuts& Lu, bool rec, float ax, float ay, float etaR, float lr, vec3 wo, vec3 wi) {
  if (!same_hemisphere(wo, wi)) return 0.f;
  float w = eval_ggx_reflection_dielectric(etaR, ax, ay, wo, wi);
  w *= fabsf(wi.y);
  vec3 wm = normalize(wo + wi);
  float fms = fms_dielectric(Lu, rec, etaR, lr, fabsf(dot(wo, wm)));
  float ms = energy_compensation(Lu, rec, fms, fabsf(wo.y), fabsf(wi.y), lr, w);
  return w + ms;
}

struct SpecularLayer {
  float ax, ay, lr, etaR;
  bool allowRefl, allowRefr;

  YPT_HD float ms_sampling_probability(const MSLuts& Lu, bool rec, vec3 wo) const {
    if (!rec) return 0.f;
    if (!allowRefr || is_delta_ggx2(ax, ay)) return 0.f;
    return saturate(1.f - directional_albedo_translucent(Lu, etaR, wo.y, lr));
  }
  YPT_HD float ms_refraction_selection_probability(const MSLuts& Lu) const {
    return allowRefl ? ms_transmission_ratio(Lu, etaR, lr) : 1.f;
  }
  YPT_HD vec3 sample_wi(const MSLuts& Lu, bool rec, vec3 wo, float s0, float s1, float sel) const {
    float msProb = ms_sampling_probability(Lu, rec, wo);
    if (sel < msProb) {
      float sc = sel / msProb;
      vec3 wi = sample_cosine_hemisphere(s0, s1);
      bool msRefract = sc < ms_refraction_selection_probability(Lu);
      wi.y = ((wo.y < 0.f) != msRefract) ? -wi.y : wi.y;
      return wi;
    }
    sel = (sel - msProb) / (1.f - msProb);
    return sample_dielectric(etaR, ax, ay, wo, s0, s1, sel, allowRefl, allowRefr);
  }
  YPT_HD float evaluate(const MSLuts& Lu, bool rec, vec3 wo, vec3 wi) const {
    float w;
    if (same_hemisphere(wo, wi)) {
      w = eval_ggx_reflection_dielectric(etaR, ax, ay, wo, wi);
      float ms;
      if (allowRefl && allowRefr) {
        ms = energy_compensation_translucent(Lu, rec, etaR, wo.y, wi.y, lr, w);
      } else {
        vec3 wm = normalize(wo + wi);
        float fms = fms_dielectric(Lu, rec, etaR, lr, fabsf(dot(wo, wm)));
        ms = energy_compensation(Lu, rec, fms, fabsf(wo.y), fabsf(wi.y), lr, w);
      }
      w += ms;
      w *= fabsf(wi.y);
    } else {
      w = eval_ggx_transmitted(etaR, ax, ay, wo, wi);
      float ms = energy_compensation_translucent(Lu, rec, etaR, wo.y, wi.y, lr, w);
      w += ms;
      w *= fabsf(wi.y);
    }
    return w;
  }
  YPT_HD float pdf(const MSLuts& Lu, bool rec, vec3 wo, vec3 wi) const {
    float pdfSS = pdf_dielectric(etaR, wo, wi, ax, ay, allowRefl, allowRefr);
    float msProb = ms_sampling_probability(Lu, rec, wo);
    if (msProb == 0.f) return pdfSS;
    float refrSel = ms_refraction_selection_probability(Lu);
    float pdfMS = pdf_cosine_hemisphere(fabsf(wi.y)) * (same_hemisphere(wo, wi) ? 1.f - refrSel : refrSel);
    return lerpf(pdfSS, pdfMS, msProb);
  }
  YPT_HD float energy_left(const MSLuts& Lu, vec3 wo, vec3 wi, float layerWeight) const {
    if (same_hemisphere(wo, wi)) return energy_remaining_after_specular(Lu, etaR, wo.y, lr, layerWeight);
    return 0.f;
  }
};

YPT_HD SpecularLayer make_specular_layer(float ax, float ay, float lr, float eta, bool refl, bool refr) {
  SpecularLayer l;
  l.ax = ax; l.ay = ay; l.lr = lr; l.etaR = eta; l.allowRefl = refl; l.allowRefr = refr;
  return l;
}

YPT_HD Spectrum conductor_eval(const MSLuts& Lu, bool rec, const Spectrum& c0, const Spectrum& c1, float ax, float ay,
                               float lr, float fromIOR, vec3 wo, vec3 wi) {
  if (!same_hemisphere(wo, wi)) return Spectrum(0.f);
  Spectrum real, img;
  for (int i = 0; i < 4; ++i) {
    float n, k2;
    conductor_n_k2(c0.s[i], c1.s[i], n, k2);
    real.s[i] = n;
    img.s[i] = sqrtf(k2);
  }
  Spectrum etaR = real / fromIOR;
  Spectrum etaK = img / fromIOR;
  Spectrum w = eval_ggx_reflection_conductor(etaR, etaK, ax, ay, wo, wi);
  vec3 wm = normalize(wo + wi);
  Spectrum fms = fresnel_conductor(etaR, etaK, dot(wo, wm));
  Spectrum ms = energy_compensation(Lu, rec, fms, wo.y, wi.y, lr, w);
  return (w + ms) * fabsf(wi.y);
}

YPT_HD float dispersed_ior(const SurfaceDef& s, const Lambdas&) { return s.dispersedIOR; }

// getSampleDirectionOSForLayer (materialSample.hlsl:148-262)
YPT_HD vec3 sample_layer_direction(const SurfaceContext& c, const MSLuts& Lu, int layer, float randSel, float s0, float s1,
                                   bool& isDiffuse, bool& isDelta) {
  const SurfaceDef& s = c.def;
  const PrecalcSurface& p = c.pre;
  vec3 wi(0.f);
  bool allowTransmitted = s.transparency > 0.f;
  float fromIOR = p.fromIOR, toIOR = p.toIOR;
  isDiffuse = false;
  isDelta = false;
  if (layer == LAYER_COATING_GGX) {
    if (p.woCoating.y > 0.f) {
      float cax, cay;
      roughness_params(s.clearCoatRoughness, 0.f, cax, cay);
      wi = sample_ggx_reflection_dielectric(cax, cay, p.woCoating, s0, s1);
      wi = mul(wi, p.to_coating(s.tangent));
      isDiffuse = considered_diffuse_ggx(cax, cay);
      isDelta = is_delta_ggx2(cax, cay);
    }
  } else if (layer == LAYER_COATING_SHEEN) {
    if (p.woBase.y > 0.f) {
      wi = sample_sheen(s.sheenRoughness, p.woBase, s0, s1);
      wi = mul(wi, p.to_base(s.tangent));
      isDiffuse = considered_diffuse_sheen(s.sheenRoughness);
      isDelta = is_delta_sheen(s.sheenRoughness);
    }
  } else if (layer == LAYER_SPEC_CONDUCTOR) {
    if (p.woBase.y > 0.f) {
      wi = sample_ggx_reflection_conductor(p.ax, p.ay, p.woBase, s0, s1);
      wi = mul(wi, p.to_base(s.tangent));
      isDiffuse = considered_diffuse_ggx(p.ax, p.ay);
      isDelta = is_delta_ggx2(p.ax, p.ay);
    }
  } else if (layer == LAYER_SPEC_DIELECTRIC) {
    if (has_dispersion(s.flags)) {
      if (p.exiting) fromIOR = dispersed_ior(s, c.lambdas);
      else toIOR = dispersed_ior(s, c.lambdas);
    }
    float etaR = toIOR / fromIOR;
    SpecularLayer sl = make_specular_layer(p.ax, p.ay, s.roughness, etaR, s.specularAmount > 0.f, s.transparency > 0.f);
    wi = sl.sample_wi(Lu, c.reciprocal, p.woBase, s0, s1, randSel);
    wi = mul(wi, p.to_base(s.tangent));
    isDiffuse = considered_diffuse_ggx(p.ax, p.ay);
    isDelta = is_delta_ggx2(p.ax, p.ay);
  } else if (layer == LAYER_DIFFUSE_REFL) {
    if (p.woBase.y > 0.f) {
      wi = sample_diffuse_lambertian(p.woBase, s0, s1);
      wi = mul(wi, p.to_base(s.tangent));
      isDiffuse = true;
      isDelta = false;
    }
  }
  if (!allowTransmitted && dot(s.geometryNormal, wi) < 0.f) wi = vec3(0.f);
  // REJECT_SHADING_GEOMETRY_FRAME_MISMATCH
  if (allowTransmitted && layer == LAYER_SPEC_DIELECTRIC && !is_zero(wi)) {
    vec3 woOS = mul(p.woBase, p.to_base(s.tangent));
    bool transmittedGeometry = dot(woOS, s.geometryNormal) * dot(wi, s.geometryNormal) < 0.f;
    bool transmittedShading = p.woBase.y * dot(wi, p.baseNormal) < 0.f;
    if (transmittedGeometry != transmittedShading) wi = vec3(0.f);
  }
  return wi;
}

YPT_HD float shading_normal_compensation(float woS, float wiS, float woG, float wiG) { return (woS * wiG) / (woG * wiS); }

// evaluateSurface (materialSample.hlsl:269-562)
YPT_HD SurfaceEval evaluate_surface(const SurfaceContext& c, const MSLuts& Lu, vec3 woObj, vec3 wiObj, uint32_t flags) {
  const SurfaceDef& s = c.def;
  const PrecalcSurface& p = c.pre;
  const bool rec = c.reciprocal;
  SurfaceEval res;
  float ax = p.ax, ay = p.ay;
  float cax, cay;
  roughness_params(s.clearCoatRoughness, 0.f, cax, cay);
  vec3 woCoating = p.woCoating, woBase = p.woBase;
  bool exiting = p.exiting;
  float fromIOR = p.fromIOR, toIOR = p.toIOR;
  float specOcc = s.specOcclusion, diffOcc = s.diffOcclusion;
  Spectrum weightSum(0.f);
  float pdfSum = 0.f;
  vec3 wiCoating = normalize(mul(p.to_coating(s.tangent), wiObj));
  vec3 wiBase = normalize(mul(p.to_base(s.tangent), wiObj));
  res.transmission = TRANSMISSION_NONE;
  bool onlyPDF = (flags & EVAL_PDF_ONLY) != 0;
  bool lightPath = (flags & EVAL_LIGHT_PATH) != 0;
  bool treatAsDelta = (flags & EVAL_TREAT_AS_DELTA) != 0;
  bool negateProjection = (flags & EVAL_DONT_PROJECT_SOLID_ANGLE) != 0;
  float dotWoG = dot(woObj, s.geometryNormal);
  float dotWiG = dot(wiObj, s.geometryNormal);
  bool baseReflected = dotWoG * dotWiG > 0.f;
  bool baseRefracted = dotWoG * dotWiG < 0.f;
  float energyLeft = 1.f;
  float energyLeftFromLayer = 1.f;
  float sncCoating = lightPath ? shading_normal_compensation(fabsf(woCoating.y), fabsf(wiCoating.y), fabsf(dotWoG), fabsf(dotWiG)) : 1.f;
  float sncBase = lightPath ? shading_normal_compensation(fabsf(woBase.y), fabsf(wiBase.y), fabsf(dotWoG), fabsf(dotWiG)) : 1.f;
  if (negateProjection) {
    sncCoating = safe_div(sncCoating, fabsf(wiCoating.y));
    sncBase = safe_div(sncBase, fabsf(wiBase.y));
  }
  float dielAmount = 1.f - s.metalness;
  float specAmount = dielAmount * s.specularAmount;
  bool applyInference = s.thinFilmThicknessNM > 0.f && s.clearCoatAmount > 0.f && (s.metalness > 0.f || specAmount > 0.f);
  applyInference = applyInference && (is_delta_ggx2(cax, cay) == is_delta_ggx2(ax, ay));
  Spectrum coatingAmplitude(0.f);

  // coating layer (ggx & sheen)
  if (s.clearCoatAmount > 0.f) {
    float etaR = s.clearCoatIOR / fromIOR;
    if (is_delta_ggx2(cax, cay) == treatAsDelta) {
      float pdf = same_hemisphere(woCoating, wiCoating) ? pdf_ggx_reflection_dielectric(woCoating, wiCoating, cax, cay) : 0.f;
      bool coatingReflected = dot(woObj, s.geometryNormal) * dot(wiObj, s.geometryNormal) > 0.f;
      if (pdf > 0.f && coatingReflected) {
        if (!onlyPDF) {
          float w = coat_eval(Lu, rec, cax, cay, etaR, s.clearCoatRoughness, woCoating, wiCoating);
          Spectrum weight = Spectrum(w) * (s.clearCoatAmount * energyLeft);
          weight = weight * sncCoating * specOcc;
          if (applyInference) coatingAmplitude = weight;
          else weightSum = weightSum + weight;
        }
        pdfSum += c.prob[LAYER_COATING_GGX] * pdf;
      }
    }
    energyLeftFromLayer *= energy_remaining_after_specular(Lu, etaR, woCoating.y, s.clearCoatRoughness, s.clearCoatAmount);
  }
  if (s.sheenAmount > 0.f) {
    if (is_delta_sheen(s.sheenRoughness) == treatAsDelta) {
      float pdf = same_hemisphere(woBase, wiBase) ? pdf_sheen(woBase, wiBase, s.sheenRoughness) : 0.f;
      if (pdf > 0.f && baseReflected) {
        if (!onlyPDF) {
          Spectrum w = same_hemisphere(woBase, wiBase) ? eval_sheen(s.sheenColor, s.sheenRoughness, woBase, wiBase) * fabsf(wiBase.y) : Spectrum(0.f);
          w = w * (s.sheenAmount * energyLeft);
          weightSum = weightSum + w * sncBase * specOcc;
        }
        pdfSum += c.prob[LAYER_COATING_SHEEN] * pdf;
      }
    }
    energyLeftFromLayer *= energy_remaining_after_sheen(Lu, woBase.y, s.sheenRoughness, s.sheenAmount);
  }
  energyLeft *= energyLeftFromLayer;
  energyLeftFromLayer = 1.f;

  // specular layers: conductor and dielectric
  if (s.metalness > 0.f) {
    if (is_delta_ggx2(ax, ay) == treatAsDelta) {
      float pdf = same_hemisphere(woBase, wiBase) ? pdf_ggx_reflection_conductor(woBase, wiBase, ax, ay) : 0.f;
      if (pdf > 0.f && baseReflected) {
        if (!onlyPDF) {
          Spectrum w = conductor_eval(Lu, rec, s.albedo, s.specular, ax, ay, s.roughness, fromIOR, woBase, wiBase);
          w = w * (s.metalness * energyLeft);
          Spectrum weight = w * sncBase * specOcc;
          if (applyInference) {
            float etaRCoating = s.clearCoatIOR / fromIOR;
            vec3 wm = normalize(woCoating + wiCoating);
            float phaseShift = etaRCoating > 1.f ? 0.f : 0.5f;
            weight = apply_thin_film(c.lambdas, coatingAmplitude, weight, s.clearCoatIOR, etaRCoating, phaseShift,
                                     dot(woCoating, wm), s.thinFilmThicknessNM);
          }
          weightSum = weightSum + weight;
        }
        pdfSum += c.prob[LAYER_SPEC_CONDUCTOR] * pdf;
      }
    }
    energyLeftFromLayer *= (1.f - s.metalness);
  }
  if (specAmount > 0.f && baseReflected) {
    if (has_dispersion(s.flags)) {
      if (exiting) fromIOR = dispersed_ior(s, c.lambdas);
      else toIOR = dispersed_ior(s, c.lambdas);
    }
    float etaR = toIOR / fromIOR;
    SpecularLayer sl = make_specular_layer(ax, ay, s.roughness, etaR, s.specularAmount > 0.f, s.transparency > 0.f);
    if (is_delta_ggx2(ax, ay) == treatAsDelta) {
      float pdf = sl.pdf(Lu, rec, woBase, wiBase);
      if (pdf > 0.f) {
        if (!onlyPDF) {
          Spectrum weight = Spectrum(sl.evaluate(Lu, rec, woBase, wiBase)) * (specAmount * energyLeft);
          weight = weight * sncBase * specOcc;
          if (applyInference) {
            float etaRCoating = s.clearCoatIOR / fromIOR;
            vec3 wm = normalize(woCoating + wiCoating);
            float phaseShift = ((etaRCoating > 1.f) != (toIOR > s.clearCoatIOR)) ? 0.5f : 0.f;
            weight = apply_thin_film(c.lambdas, coatingAmplitude, weight, s.clearCoatIOR, etaRCoating, phaseShift,
                                     dot(woCoating, wm), s.thinFilmThicknessNM);
          }
          weightSum = weightSum + weight;
        }
        pdfSum += c.prob[LAYER_SPEC_DIELECTRIC] * pdf;
      }
    }
    energyLeftFromLayer *= sl.energy_left(Lu, woBase, wiBase, specAmount);
  }
  energyLeft *= energyLeftFromLayer;
  energyLeftFromLayer = 1.f;

  // bottom layer: diffuse and transmission
  float diffuseAmount = (1.f - s.transparency) * dielAmount * (1.f - s.subsurfaceAmount);
  if (diffuseAmount > 0.f && !treatAsDelta) {
    float pdf = same_hemisphere(woBase, wiBase) ? pdf_diffuse_lambertian(wiBase) : 0.f;
    if (pdf > 0.f && baseReflected) {
      if (!onlyPDF) {
        Spectrum w = same_hemisphere(woBase, wiBase) ? (s.albedo / PI) * fabsf(wiBase.y) : Spectrum(0.f);
        w = w * (diffuseAmount * energyLeft);
        weightSum = weightSum + w * sncBase * diffOcc;
        energyLeftFromLayer *= 1.f - diffuseAmount;
      }
      pdfSum += c.prob[LAYER_DIFFUSE_REFL] * pdf;
    }
  }
  float transparencyAmount = s.transparency * dielAmount;
  if (transparencyAmount > 0.f && baseRefracted && (is_delta_ggx2(ax, ay) == treatAsDelta)) {
    if (has_dispersion(s.flags)) {
      if (exiting) fromIOR = dispersed_ior(s, c.lambdas);
      else toIOR = dispersed_ior(s, c.lambdas);
      res.transmission |= TRANSMISSION_DISPERSED;
    }
    float etaR = toIOR / fromIOR;
    SpecularLayer sl = make_specular_layer(ax, ay, s.roughness, etaR, s.specularAmount > 0.f, s.transparency > 0.f);
    float pdf = sl.pdf(Lu, rec, woBase, wiBase);
    if (pdf > 0.f) {
      pdfSum += c.prob[LAYER_SPEC_DIELECTRIC] * pdf;
      bool twoSided = surface_two_sided(s.flags);
      bool wasTransmitted = !twoSided;
      float solidAngleCompression = 1.f;
      if (wasTransmitted) {
        if (!lightPath) solidAngleCompression *= sqr(1.f / etaR);
        res.transmission |= exiting ? TRANSMISSION_EXITED : TRANSMISSION_ENTERED;
      } else {
        res.transmission |= TRANSMISSION_ENTERED | TRANSMISSION_EXITED;
      }
      if (!onlyPDF) {
        Spectrum w = Spectrum(sl.evaluate(Lu, rec, woBase, wiBase)) * (transparencyAmount * energyLeft);
        weightSum = weightSum + w * solidAngleCompression * sncBase * specOcc;
      }
    }
  }
  res.pdf = pdfSum;
  res.weight = weightSum;
  return res;
}

}  // namespace ypt
