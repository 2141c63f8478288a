// Per-vertex work of the unidirectional integrators (BPT and BPT_WITH_NEE).
//
// Restates shaders/raytracing/rayGenPrimaryRays.hlsl:10-145 (path start, bounce loop, Russian
// roulette), rayHitDefault.hlsl:27-585 (explicit light connection, MIS, surface evaluation and next
// direction), rayMissEnvironment.hlsl:13-31, hitShadersCommon.hlsl:138-160 (transmittance, ray
// origin offset), lightSampling.hlsl:31-228,360-370 (light pick, light sampling, density conversion),
// raytraceCommonResources.hlsl:108-156 (thin-lens camera) and payload.hlsl (volume stack).
//
// The shadow ray of an explicit light connection is returned to the caller instead of being traced
// here, so that one driver can trace it at once and another can batch all of them in a separate pass:
// the contribution and the order in which it reaches totalLight are the same either way.
#pragma once
#include "scene.h"
#include "sampler.h"
#include "subsurface.h"

namespace ypt {

static constexpr uint32_t RAY_ALIVE = 0, RAY_TERMINATED = 1, RAY_CANCELLED = 2;
static constexpr uint32_t RS_SECONDARY_TERMINATED = 1u << 0;
static constexpr uint32_t RS_REGULARIZE = 1u << 1;
static constexpr uint32_t RS_DELTA = 1u << 2;
static constexpr uint32_t VOLUME_NONE = 0xFFFFFFFFu;
static constexpr int MAX_VOLUMES = 4;
static constexpr uint32_t NO_HIT = 0xFFFFFFFFu;

struct PathState {
  Spectrum throughput, totalLight;
  vec3 origin, dir;
  float lambda;
  float pdfThisRay;
  uint32_t dim, seed, flags, pathLength, rayState, nvol;
  uint32_t furnaceInstance;  // WHITE_FURNACE_TEST: first object the path hit (rayIndex in the reference)
  float volIor[MAX_VOLUMES];
  Spectrum volAbs[MAX_VOLUMES];
  uint32_t volMat[MAX_VOLUMES];
};

struct ShadowRay {
  vec3 origin, dir;
  float tmax;
  uint32_t targetTri;  // closest hit must be this triangle for the light to count; NO_HIT: nothing may be hit
  Spectrum contrib;    // added to totalLight when visible
  bool valid;
};

// first-hit features (hitShadersCommon.hlsl:94-136), filled when requested
struct FirstHitFeatures {
  vec3 normal;
  vec3 albedo;
  float roughness, ior, metalness, transparency, anisotropy, depth;
  uint32_t hit;
};

// Per-sample feature record in device memory: normal xyz, albedo rgb, roughness, ior, metalness,
// transparency, depth, hit (1 or 0), anisotropy, 3 unused; all zero for a miss. The first 12 floats
// are what ypt_get_first_hit_features returns per pixel.
static constexpr uint32_t FEATURE_STRIDE = 16;
YPT_HD void store_features(float* f, const FirstHitFeatures& fh) {
  if (!fh.hit) {
    for (uint32_t k = 0; k < FEATURE_STRIDE; ++k) f[k] = 0.f;
    return;
  }
  f[0] = fh.normal.x; f[1] = fh.normal.y; f[2] = fh.normal.z;
  f[3] = fh.albedo.x; f[4] = fh.albedo.y; f[5] = fh.albedo.z;
  f[6] = fh.roughness; f[7] = fh.ior; f[8] = fh.metalness; f[9] = fh.transparency; f[10] = fh.depth; f[11] = 1.f;
  f[12] = fh.anisotropy; f[13] = 0.f; f[14] = 0.f; f[15] = 0.f;
}

// ---------------------------------------------------------------- volume stack (payload.hlsl)
YPT_HD float current_ior(const PathState& p) { return p.nvol != 0 ? sel_get(p.volIor, p.nvol - 1) : 1.0f; }
YPT_HD float previous_ior(const PathState& p) { return p.nvol > 1 ? sel_get(p.volIor, p.nvol - 2) : 1.0f; }
YPT_HD Spectrum current_absorption(const PathState& p) { return p.nvol != 0 ? sel_get(p.volAbs, p.nvol - 1) : Spectrum(0.f); }
YPT_HD void entered_volume(PathState& p, float ior, const Spectrum& absorption, uint32_t mat) {
  p.nvol = p.nvol + 1 < (uint32_t)MAX_VOLUMES ? p.nvol + 1 : (uint32_t)MAX_VOLUMES;
  sel_set(p.volIor, p.nvol - 1, ior);
  sel_set(p.volAbs, p.nvol - 1, absorption);
  sel_set(p.volMat, p.nvol - 1, mat);
}
// removes the innermost entry of material mat (the innermost entry when mat is not on the stack)
YPT_HD void exited_volume(PathState& p, uint32_t mat) {
  if (p.nvol == 0) return;
  uint32_t index = p.nvol - 1;
#pragma unroll
  for (int i = 0; i < MAX_VOLUMES; ++i) index = ((uint32_t)i < p.nvol && p.volMat[i] == mat) ? (uint32_t)i : index;
#pragma unroll
  for (int i = 0; i + 1 < MAX_VOLUMES; ++i) {
    bool shift = (uint32_t)i >= index && (uint32_t)i + 1 < p.nvol;
    p.volIor[i] = pick(shift, p.volIor[i + 1], p.volIor[i]);
    p.volAbs[i] = pick(shift, p.volAbs[i + 1], p.volAbs[i]);
    p.volMat[i] = pick(shift, p.volMat[i + 1], p.volMat[i]);
  }
  --p.nvol;
}

// "A Fast and Robust Method for Avoiding Self-Intersection", Ray Tracing Gems ch. 6 (hitShadersCommon.hlsl:147-160)
YPT_HD vec3 spawn_ray_origin(vec3 position, vec3 geometryNormal, vec3 rayDir) {
  const float originThreshold = 1.0f / 32.0f;
  const float floatScale = 1.0f / 65536.0f;
  const float intScale = 256.0f;
  vec3 n = dot(rayDir, geometryNormal) < 0.f ? -geometryNormal : geometryNormal;
  int ox = (int)(intScale * n.x), oy = (int)(intScale * n.y), oz = (int)(intScale * n.z);
  int sx = position.x < 0.f ? -ox : ox;
  int sy = position.y < 0.f ? -oy : oy;
  int sz = position.z < 0.f ? -oz : oz;
  float px = as_float_i(as_int(position.x) + sx);
  float py = as_float_i(as_int(position.y) + sy);
  float pz = as_float_i(as_int(position.z) + sz);
  return vec3(fabsf(position.x) < originThreshold ? position.x + floatScale * n.x : px,
              fabsf(position.y) < originThreshold ? position.y + floatScale * n.y : py,
              fabsf(position.z) < originThreshold ? position.z + floatScale * n.z : pz);
}

YPT_HD float weight_mis(float a, float b) {  // power heuristic
  float a2 = a * a, b2 = b * b;
  return a2 / (a2 + b2);
}
YPT_HD float solid_angle_to_area(vec3 v, vec3 n) {
  float invDistSqr = 1.f / dot(v, v);
  float absDot = fabsf(dot(n, v * sqrtf(invDistSqr)));
  return absDot * invDistSqr;
}
YPT_HD float area_to_solid_angle(vec3 v, vec3 n) { return 1.f / solid_angle_to_area(v, n); }

// ---------------------------------------------------------------- camera
YPT_HD vec3 xform_point(const mat4& M, vec3 p) {
  float v[4] = {p.x, p.y, p.z, 1.f}, o[4];
  mul4(M, v, o);
  return vec3(o[0], o[1], o[2]);
}
YPT_HD vec3 xform_dir(const mat4& M, vec3 d) {
  float v[4] = {d.x, d.y, d.z, 0.f}, o[4];
  mul4(M, v, o);
  return vec3(o[0], o[1], o[2]);
}

// generateCameraRay (raytraceCommonResources.hlsl:115-137): thin-lens ray through texel (px, py),
// texel jitter (jx, jy) and lens sample (l0, l1)
YPT_HD void camera_ray(const FrameConstants& fc, uint32_t px, uint32_t py, float jx, float jy, float l0, float l1,
                       vec3& origin, vec3& dir) {
  float uvx = (float)px * fc.targetTexDimensions[2] + fc.rayUVOffset[0];
  float uvy = (float)py * fc.targetTexDimensions[3] + fc.rayUVOffset[1];
  float ox = 0.f, oy = 0.f;
  if ((fc.flags & FRAME_FLAG_DISABLE_JITTER) == 0) { ox = jx - 0.5f; oy = jy - 0.5f; }
  ox *= fc.targetTexDimensions[2];
  oy *= fc.targetTexDimensions[3];
  float nearp[4] = {uvx + ox, uvy + oy, 0.0f, 1.0f}, v[4];
  mul4(fc.uvToView, nearp, v);
  vec3 pinhole(v[0] / v[3], v[1] / v[3], v[2] / v[3]);
  vec3 focusPoint = pinhole * safe_div(fc.lensParams[1], fabsf(pinhole.z));
  float dx, dy;
  sample_concentric_disk(l0, l1, dx, dy);
  vec3 lensPos(dx * fc.lensParams[0], dy * fc.lensParams[0], 0.f);
  origin = xform_point(fc.viewToWorld, lensPos);
  dir = normalize(xform_dir(fc.viewToWorld, focusPoint - lensPos));
}

// rayGenPrimaryRays.hlsl:13-39
YPT_HD void start_path(const FrameConstants& fc, const uint32_t* point, uint32_t px, uint32_t py, PathState& ps) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0];
  uint32_t rayIndex = py * W + px;
  RandomSampler rand;
  rand.init(point, 0u, rayIndex);
  float lambda = hero_wavelength(rand.r1());
  float jx, jy;
  rand.texel_offset(jx, jy);
  float lens[2];
  rand.r2(lens);
  camera_ray(fc, px, py, jx, jy, lens[0], lens[1], ps.origin, ps.dir);
  ps.throughput.set(1.f);
  ps.totalLight.set(0.f);
  ps.rayState = RAY_ALIVE;
  ps.pathLength = 0;
  ps.nvol = 0;
  for (int i = 0; i < MAX_VOLUMES; ++i) { ps.volMat[i] = VOLUME_NONE; ps.volIor[i] = 1.f; ps.volAbs[i].set(0.f); }
  ps.flags = 0;
  ps.furnaceInstance = VOLUME_NONE;
  ps.pdfThisRay = 0.f;
  ps.dim = rand.dim;
  ps.seed = rand.seed;
  ps.lambda = lambda;
}

// ---------------------------------------------------------------- miss
YPT_HD Spectrum rgb_unbounded_spectrum(const SceneView& S, vec3 rgb, const Lambdas& L);

YPT_HD void shade_miss(const SceneView& S, const FrameConstants& fc, PathState& ps) {
  Lambdas L = spectral_lambdas(ps.lambda);
  vec3 color;
  if (fc.flags & FRAME_FLAG_WHITE_FURNACE) color = vec3(1.f);
  else color = env_color(S, ps.dir) * fc.envIntensityScale;  // getSkyBoxColor (rayMissEnvironment.hlsl:23)
  ps.totalLight = ps.totalLight + ps.throughput * rgb_unbounded_spectrum(S, color, L);
  ps.rayState = RAY_TERMINATED;
}

// RGB -> coefficients via the 64^3 table (only needed for non-grey colours on the device).
YPT_HD void rgb_coeffs(const SceneView& S, vec3 v, float c[3]) { rgb_coeffs_table(S.rgb2spec, v, c); }

YPT_HD Spectrum rgb_unbounded_spectrum(const SceneView& S, vec3 values, const Lambdas& L) {  // setFromRGBUnbounded
  float m = fmax_(fmax_(values.x, values.y), values.z);
  float scale = 2.f * m;
  vec3 nv = scale == 0.f ? vec3(0.f) : values / scale;
  SpectralColor col;
  rgb_coeffs(S, nv, col.c);
  col.scale = scale;
  return spectrum_unbounded(col, L, S.spec);
}

// ---------------------------------------------------------------- lights
// sampleLightIndex / lightSelectionPDF (lightSampling.hlsl:31-42)
YPT_HD uint32_t sample_light_index(const LightGPU* lights, uint32_t count, float u, float& pickPdf) {
  float scaled = u * (float)count;
  uint32_t index = (uint32_t)scaled;
  if (index > count - 1) index = count - 1;
  const LightGPU& e = lights[index];
  uint32_t chosen = (scaled - (float)index) < e.aliasThreshold ? index : e.aliasIndex;
  pickPdf = lights[chosen].selectionPdf;
  return chosen;
}

// emission of material m at a point of triangle tri (fetchSurfaceMaterialParameters +
// modifySurfaceEmissionWithTexture, materialsCommon.hlsl:420-428)
YPT_HD Spectrum surface_emission(const SceneView& S, const MaterialGPU& m, uint32_t tri, float b1, float b2, const Lambdas& L) {
  if (m.tex[TEX_EMISSIVE] < 0) {
    if (m.blackbodyK > 0.f) return blackbody_spectrum(L, m.blackbodyK, fmax_(fmax_(m.emissiveRGB[0], m.emissiveRGB[1]), m.emissiveRGB[2]));
    return spectrum_unbounded(m.emissive, L, S.spec);
  }
  float u, v, e[4];
  surface_uv(S, tri, b1, b2, u, v);
  sample_texture(S, (uint32_t)m.tex[TEX_EMISSIVE], u * m.texScale[2 * TEX_EMISSIVE], v * m.texScale[2 * TEX_EMISSIVE + 1], e);
  vec3 rgb(m.emissiveRGB[0] * e[0], m.emissiveRGB[1] * e[1], m.emissiveRGB[2] * e[2]);
  if (m.blackbodyK > 0.f) return blackbody_spectrum(L, m.blackbodyK, fmax_(fmax_(rgb.x, rgb.y), rgb.z));
  return rgb_unbounded_spectrum(S, rgb, L);
}

// samplePunctualLightIndex / punctualLightSelectionPDF (lightSampling.hlsl:44-57): the punctual
// lights' own alias table
YPT_HD uint32_t sample_punctual_light_index(const LightGPU* lights, uint32_t count, uint32_t punctualCount, float u,
                                            float& pickPdf) {
  uint32_t first = count - punctualCount;
  float scaled = u * (float)punctualCount;
  uint32_t index = (uint32_t)scaled;
  if (index > punctualCount - 1) index = punctualCount - 1;
  const LightGPU& e = lights[first + index];
  uint32_t chosen = first + ((scaled - (float)index) < e.punctualAliasThreshold ? index : e.punctualAliasIndex);
  pickPdf = lights[chosen].punctualSelectionPdf;
  return chosen;
}

YPT_HD float smoothstep_(float a, float b, float x) {  // HLSL smoothstep
  float t = saturate((x - a) / (b - a));
  return t * t * (3.f - 2.f * t);
}

// spotFalloff / punctualLightAngularFalloff (lightSampling.hlsl:238-256)
YPT_HD float punctual_falloff(const LightGPU& l, vec3 dirFromLight) {
  if (l.type != LIGHT_TYPE_SPOT) return 1.f;
  vec3 axis(l.directionCosOuter[0], l.directionCosOuter[1], l.directionCosOuter[2]);
  float cosTheta = dot(dirFromLight, axis);
  float cosInner = l.positionCosInner[3], cosOuter = l.directionCosOuter[3];
  if (cosInner <= cosOuter) return cosTheta >= cosOuter ? 1.f : 0.f;
  return smoothstep_(cosOuter, cosInner, cosTheta);
}

// evaluatePunctualLightIncidence (lightSampling.hlsl:258-284): direction and distance to the light
// (infinite for directional lights) and the incident RGB intensity at pos
YPT_HD vec3 punctual_incidence(const LightGPU& l, vec3 pos, vec3& toLight, float& distance) {
  vec3 I(l.intensity[0], l.intensity[1], l.intensity[2]);
  if (l.type == LIGHT_TYPE_DIRECTIONAL) {
    toLight = -vec3(l.directionCosOuter[0], l.directionCosOuter[1], l.directionCosOuter[2]);
    distance = __builtin_inff();
    return I;
  }
  vec3 d = vec3(l.positionCosInner[0], l.positionCosInner[1], l.positionCosInner[2]) - pos;
  float distSqr = fmax_(dot(d, d), 1e-10f);
  distance = sqrtf(distSqr);
  toLight = d / distance;
  return I * (punctual_falloff(l, -toLight) / distSqr);
}

struct ExplicitLightSample {
  Spectrum emission;  // radiance with emission profile (mesh lights), incident intensity (punctual)
  vec3 toLight;
  float distance;     // infinite for directional lights
  float pdf;          // solid-angle pdf with the pick (mesh lights), the pick alone (punctual)
  uint32_t tri;       // NO_HIT for punctual lights
  bool punctual;
};

// sampleExplicitLight (rayHitDefault.hlsl:36-82) with sampleLight (lightSampling.hlsl:134-197): NEE picks
// from all lights, BPT from the punctual lights only (BSDF sampling never reaches them)
YPT_HD void sample_explicit_light(const SceneView& S, const FrameConstants& fc, bool nee, vec3 pos, const float r[4],
                                  const Lambdas& L, ExplicitLightSample& out) {
  float pickPdf;
  uint32_t li = nee ? sample_light_index(S.lights, fc.lightCount, r[3], pickPdf)
                    : sample_punctual_light_index(S.lights, fc.lightCount, fc.punctualLightCount, r[3], pickPdf);
  const LightGPU& light = S.lights[li];
  out.punctual = light.type != LIGHT_TYPE_MESH;
  if (out.punctual) {
    vec3 rgb = punctual_incidence(light, pos, out.toLight, out.distance);
    out.emission = rgb_unbounded_spectrum(S, rgb, L);
    out.pdf = pickPdf;
    out.tri = NO_HIT;
    return;
  }
  uint32_t primCount = light.triCount;
  uint32_t prim = (uint32_t)fmin_((float)primCount * r[2], (float)(primCount - 1u));
  float bx = r[0], by = r[1];
  if (bx + by > 1.f) { bx = 1.f - bx; by = 1.f - by; }
  float b0 = 1.f - bx - by;
  uint32_t tri = light.firstTri + prim;
  vec3 p1, p2, p3;
  tri_vertices(S, tri, p1, p2, p3);
  vec3 gn = cross(p2 - p1, p3 - p1);
  float gl = length(gn);
  gn = gn / gl;
  float area = gl * 0.5f;
  vec3 lp = b0 * p1 + bx * p2 + by * p3;
  const MaterialGPU& m = S.mats[light.material];
  bool twoSided = surface_two_sided(m.flags);
  Spectrum radiance = surface_emission(S, m, tri, bx, by, L);
  float pdfPos = 1.f / ((float)primCount * area);
  vec3 toLight = lp - pos;
  float len = length(toLight);
  float cosEmit = -dot(gn, toLight / len);
  out.emission = radiance * emission_profile(cosEmit, m.emissionFocus);
  out.pdf = pickPdf * pdfPos * area_to_solid_angle(toLight, gn);
  out.toLight = toLight / len;
  out.distance = len;
  out.tri = tri;
  if (!twoSided && dot(toLight, gn) >= 0.f) { out.emission.set(0.f); out.pdf = 0.f; }
}

// calculateExplicitLightConnectionPDF (rayHitDefault.hlsl:139-169)
YPT_HD float light_connection_pdf(const SceneView& S, vec3 fromToLight, uint32_t tri) {
  const InstanceGPU& in = S.inst[S.triInst[tri]];
  float pick = in.lightSelectionPdf;
  vec3 p1, p2, p3;
  tri_vertices(S, tri, p1, p2, p3);
  vec3 gn = cross(p2 - p1, p3 - p1);
  float gl = length(gn);
  gn = gn / gl;
  float area = gl * 0.5f;
  float posPdf = 1.f / ((float)in.triCount * area);
  return posPdf * pick * area_to_solid_angle(fromToLight, gn);
}

// writeMaterialParamsForFirstBounce (hitShadersCommon.hlsl:94-136), unpacked
YPT_HD void first_hit_features(const FrameConstants& fc, const MaterialGPU& m, const SurfaceDef& sd, vec3 hitPos,
                               FirstHitFeatures& f, const SurfaceTexels* tx = nullptr) {
  float cw = sd.clearCoatAmount, sw = sd.sheenAmount, bw = sd.specularAmount;
  float wsum = cw + sw + bw;
  if (wsum == 0.f) wsum = 1.f;
  cw /= wsum; sw /= wsum; bw /= wsum;
  vec3 n = sd.baseNormal;
  if (sd.clearCoatAmount > 0.f) n = n + sd.coatingNormal;
  f.normal = normalize(n);
  vec3 em(m.emissiveRGB[0], m.emissiveRGB[1], m.emissiveRGB[2]);
  vec3 al(m.albedoRGB[0], m.albedoRGB[1], m.albedoRGB[2]);
  if (tx && (tx->mask & (1u << TEX_ALBEDO)))
    al = vec3(al.x * tx->v[TEX_ALBEDO][0], al.y * tx->v[TEX_ALBEDO][1], al.z * tx->v[TEX_ALBEDO][2]);
  if (tx && (tx->mask & (1u << TEX_EMISSIVE)))
    em = vec3(em.x * tx->v[TEX_EMISSIVE][0], em.y * tx->v[TEX_EMISSIVE][1], em.z * tx->v[TEX_EMISSIVE][2]);
  vec3 sh(m.sheenRGB[0], m.sheenRGB[1], m.sheenRGB[2]);
  f.albedo = em / (vec3(1.f) + em) + al + sh * sd.sheenAmount;
  f.roughness = sd.roughness * bw + sd.clearCoatRoughness * cw + sd.sheenRoughness * sw;
  f.ior = sd.dielectricIOR * sd.specularAmount + sd.clearCoatIOR * sd.clearCoatAmount;
  f.anisotropy = sd.anisotropy;
  f.metalness = sd.metalness;
  vec3 vp = xform_point(fc.worldToView, hitPos);
  f.depth = vp.z * fc.lensParams[2];
  f.transparency = sd.transparency;
  f.hit = 1u;
}

// ---------------------------------------------------------------- closest hit
// rayHitDefault (rayHitDefault.hlsl:464-585) with evaluateSurfaceAndGenerateNextSampleDirection (:316-435)
// and connectToExplicitLight (:204-257). `nee` selects BPT_WITH_NEE.
// Subsurface policy of shade_hit: NoSubsurface ends a path that takes the subsurface layer (the
// form for scenes without subsurface materials); a policy with enabled = true brings the probe
// tracer (closest_in_range) and continues the path from the exit point.
struct NoSubsurface {
  static constexpr bool enabled = false;
  YPT_HD bool closest_in_range(vec3, vec3, float, float, uint32_t, uint32_t, Hit&) { return false; }
};

YPT_HD void continue_from_subsurface_exit(const SceneView& S, const FrameConstants& fc, bool nee, bool hasLights,
                                          PathState& ps, const MaterialGPU& m, const SubsurfaceExit& ex,
                                          RandomSampler& rand, const Lambdas& L, ShadowRay& sr);

// sr: the shadow ray of the vertex's light connection; sr2: that of the subsurface exit point
template <class SSS = NoSubsurface>
YPT_HD void shade_hit(const SceneView& S, const FrameConstants& fc, const uint32_t* point, bool nee, PathState& ps,
                      const Hit& h, ShadowRay& sr, FirstHitFeatures* fh, SSS* sss = nullptr, ShadowRay* sr2 = nullptr) {
  sr.valid = false;
  Lambdas L = spectral_lambdas(ps.lambda);
  uint32_t instId = S.triInst[h.tri];
  uint32_t i0 = S.idx[3 * h.tri], i1 = S.idx[3 * h.tri + 1], i2 = S.idx[3 * h.tri + 2];
  const InstanceGPU& inst = S.inst[instId];
  uint32_t matId = inst.material;
  const MaterialGPU& m = S.mats[matId];
  // The material's texture slots are read here, before the vertex data: the texel fetch below is
  // their first use, and a GPU waits for loads in the order it issued them, so a slot load issued
  // after the vertex loads would have its wait cover all of them.
  const bool textured = (m.tex[0] & m.tex[1] & m.tex[2] & m.tex[3]) >= 0;
  float b0 = 1.f - h.b1 - h.b2;
  // the vertex data of the hit, read together (positions, normals and, when the mesh has them,
  // tangents) before any of it is used, so that their loads are in flight at once
  vec3 p1 = load3(S.pos, i0), p2 = load3(S.pos, i1), p3 = load3(S.pos, i2);
  vec3 n0 = load3(S.nrm, i0), n1 = load3(S.nrm, i1), n2 = load3(S.nrm, i2);
  float tv[3][4] = {};
  if (S.tan) {
    const uint32_t vi[3] = {i0, i1, i2};
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 4; ++c) tv[k][c] = S.tan[4 * vi[k] + c];
  }
  vec3 geometryNormal = normalize(cross(p2 - p1, p3 - p1));
  vec3 normal = normalize(b0 * n0 + h.b1 * n1 + h.b2 * n2);
  vec3 tangent(1.f, 0.f, 0.f);  // meshHasValidTangents ? fetchMeshTangent : (1, 0, 0)
  if (S.tan) {
    vec3 a(tv[0][0] * tv[0][3], tv[0][1] * tv[0][3], tv[0][2] * tv[0][3]);
    vec3 b(tv[1][0] * tv[1][3], tv[1][1] * tv[1][3], tv[1][2] * tv[1][3]);
    vec3 c(tv[2][0] * tv[2][3], tv[2][1] * tv[2][3], tv[2][2] * tv[2][3]);
    tangent = normalize(b0 * a + h.b1 * b + h.b2 * c);
  }
  vec3 rayDir = normalize(ps.dir);
  bool frontFace = dot(ps.dir, cross(p2 - p1, p3 - p1)) < 0.f;

  if (ps.pathLength >= fc.regularizeAfterVertices) ps.flags |= RS_REGULARIZE;
  bool regularizePath = (ps.flags & RS_REGULARIZE) != 0;
  SurfaceTexels tx;
  tx.mask = 0;
  tx.rgb2spec = S.rgb2spec;
  if (textured) sample_surface_texels(S, m, h.tri, h.b1, h.b2, tx);
  SurfaceDef sd = make_surface(m, geometryNormal, normal, tangent, -rayDir, true, regularizePath, L, S.spec, &tx);
  vec3 hitPos = ps.origin + ps.dir * h.t;

  if (ps.pathLength == 0 && fh) first_hit_features(fc, m, sd, hitPos, *fh, &tx);

  // handleRayStartInsideMedium (hitShadersCommon.hlsl:213-222)
  if (dot(rayDir, geometryNormal) > 0.f && ps.nvol == 0 && !surface_two_sided(sd.flags))
    entered_volume(ps, sd.dielectricIOR, sd.absorption, matId);
  Spectrum absorb = current_absorption(ps);
  if (!absorb.all_equal(0.f)) {
    Spectrum tr;
    for (int i = 0; i < 4; ++i) tr.s[i] = yexp(-absorb.s[i] * h.t);
    ps.throughput = ps.throughput * tr;
  }

  if (!sd.emissive.all_equal(0.f)) {
    float wMIS = 1.f;
    if (nee && fc.lightCount > 0 && ps.pdfThisRay > 0.f && (ps.flags & RS_DELTA) == 0) {
      float lightPdf = light_connection_pdf(S, hitPos - ps.origin, h.tri);
      wMIS = weight_mis(ps.pdfThisRay, lightPdf);
    }
    float side = dot(sd.geometryNormal, -rayDir) < 0.f ? 0.f : 1.f;
    float profile = 1.f;
    if (sd.emissionFocus > 0.f) {
      vec3 gnWS = normalize(sd.geometryNormal);
      profile = emission_profile(dot(gnWS, -normalize(ps.dir)), sd.emissionFocus);
    }
    ps.totalLight = ps.totalLight + ps.throughput * sd.emissive * wMIS * side * profile;
    ps.rayState = RAY_TERMINATED;
    return;
  }

  if (fc.flags & FRAME_FLAG_WHITE_FURNACE) {  // handleWhiteFurnaceTest (rayHitDefault.hlsl:437-459)
    if (ps.furnaceInstance == VOLUME_NONE) {
      ps.furnaceInstance = instId;
    } else if (ps.furnaceInstance != instId) {
      Spectrum white = rgb_unbounded_spectrum(S, vec3(1.f), L);
      ps.totalLight = ps.totalLight + ps.throughput * white;
      ps.rayState = RAY_TERMINATED;
      return;
    }
  }

  RandomSampler rand;
  rand.point = point; rand.dim = ps.dim; rand.seed = ps.seed;
  float rb[4];
  rand.r4(rb);
  vec3 wo = -rayDir;
  SurfaceContext ctx;
  make_surface_context(ctx, sd, current_ior(ps), previous_ior(ps), wo, frontFace, L, false);
  vec3 normalWS = normalize(sd.geometryNormal);

  bool hasLights = nee ? fc.lightCount > 0 : fc.punctualLightCount > 0;  // hasExplicitlySampledLights
  if (hasLights && !is_delta_distribution(ctx.pre.ax, ctx.pre.ay, sd)) {
    float rl[4];
    rand.r4(rl);
    ExplicitLightSample ls;
    sample_explicit_light(S, fc, nee, hitPos, rl, L, ls);
    if (!ls.emission.all_equal(0.f) && ls.pdf > 0.f) {
      vec3 toLightOS = normalize(ls.toLight);
      SurfaceEval ev = evaluate_surface(ctx, S.luts, wo, toLightOS, EVAL_NONE);
      Spectrum w = ev.weight;
      if ((ev.transmission & TRANSMISSION_DISPERSED) != 0 && (ps.flags & RS_SECONDARY_TERMINATED) == 0) w.terminate_secondary();
      if (!w.all_equal(0.f)) {
        sr.origin = spawn_ray_origin(hitPos, normalWS, ls.toLight);
        sr.dir = ls.toLight;
        // punctual lights: visible when nothing is hit on the way (checkUnoccluded); delta lights are
        // reached only by explicit sampling, so there is nothing to weight against
        sr.tmax = ls.punctual ? ls.distance - RAY_MIN_T : ls.distance + RAY_MIN_T;
        sr.targetTri = ls.tri;
        float wMIS = ls.punctual ? 1.f : weight_mis(ls.pdf, ev.pdf);
        w = (w / ls.pdf) * wMIS;
        sr.contrib = ps.throughput * w * ls.emission;
        sr.valid = true;
      }
    }
  }

  int layer = select_layer(rb[0], ctx.prob);
  if constexpr (SSS::enabled) {
    if (layer == LAYER_SUBSURFACE) {  // evaluateSurfaceAndGenerateNextSampleDirection (rayHitDefault.hlsl:339-355)
      SurfacePoint entry;
      entry.pos = hitPos;
      entry.n = orient_to_same_side(-rayDir, normalWS);
      entry.inst = instId;
      entry.tri = h.tri;
      entry.b1 = h.b1;
      entry.b2 = h.b2;
      entry.frontFace = frontFace;
      Spectrum radius = spectrum_unbounded_no_illum(m.subsurfaceRadius, L);
      SubsurfaceExit ex;
      if (!sample_subsurface_exit(S, sd.albedo, radius, entry, rb[1], rb[2], rb[3], *sss, ex)) {
        ps.pdfThisRay = 0.f;
        ps.flags &= ~RS_DELTA;
        ps.throughput.set(0.f);
        ps.dim = rand.dim;
        ps.rayState = RAY_CANCELLED;
        return;
      }
      // evaluateSubsurfaceEntry: the entry weight over the density of the subsurface edge
      Spectrum entryWeight = ex.weight * (subsurface_entry_factor(ctx, S.luts) * ex.pdfPosition);
      float pdfForward = ctx.prob[LAYER_SUBSURFACE] * ex.pdfPosition;
      ps.throughput = ps.throughput * entryWeight * safe_div(1.f, pdfForward);
      continue_from_subsurface_exit(S, fc, nee, hasLights, ps, m, ex, rand, L, *sr2);
      return;
    }
  }
  bool isDiffuse = true, isDelta = false;
  vec3 wi(0.f);
  if (layer != LAYER_SUBSURFACE) wi = sample_layer_direction(ctx, S.luts, layer, rb[1], rb[2], rb[3], isDiffuse, isDelta);
  bool local = !is_zero(wi) && layer != LAYER_SUBSURFACE;
  SurfaceEval ev;
  ev.weight.set(0.f);
  ev.pdf = 0.f;
  ev.transmission = TRANSMISSION_NONE;
  if (local) {
    ev = evaluate_surface(ctx, S.luts, wo, wi, isDelta ? EVAL_TREAT_AS_DELTA : EVAL_NONE);
    if (isDiffuse) ps.flags |= RS_REGULARIZE;
  }
  ps.pdfThisRay = ev.pdf;
  if (isDelta) ps.flags |= RS_DELTA;
  else ps.flags &= ~RS_DELTA;
  if (ev.pdf > 0.f) {
    if (ev.transmission != TRANSMISSION_NONE) {
      if ((ev.transmission & TRANSMISSION_DISPERSED) != 0) {
        if ((ps.flags & RS_SECONDARY_TERMINATED) == 0) ev.weight.terminate_secondary();
        ps.flags |= RS_SECONDARY_TERMINATED;
      }
      bool entered = (ev.transmission & TRANSMISSION_ENTERED) != 0;
      bool exited = (ev.transmission & TRANSMISSION_EXITED) != 0;
      if (entered != exited) {
        if (exited) exited_volume(ps, matId);
        else entered_volume(ps, sd.dielectricIOR, sd.absorption, matId);
      }
    }
    ev.weight = ev.weight / ev.pdf;
  }
  vec3 nextDir(0.f), nextOrigin = hitPos;
  if (!is_zero(wi)) {
    vec3 nWorld = normalize(sd.geometryNormal);
    nextDir = normalize(wi);
    nextOrigin = spawn_ray_origin(hitPos, nWorld, nextDir);
  }
  ps.throughput = ps.throughput * ev.weight;
  ps.dim = rand.dim;
  if (is_zero(nextDir) || ps.pdfThisRay == 0.0f) {
    ps.rayState = RAY_CANCELLED;
  } else {
    ps.dir = nextDir;
    ps.origin = nextOrigin;
  }
}

// continueFromSubsurfaceExit (rayHitDefault.hlsl:259-314): the surface at the exit point, set up for
// the outside (the outward normal stands in for the incoming direction), a light connection through
// the exit lobe only, and the next direction from the exit lobe. The entry weight is already in
// the throughput.
YPT_HD void continue_from_subsurface_exit(const SceneView& S, const FrameConstants& fc, bool nee, bool hasLights,
                                          PathState& ps, const MaterialGPU& m, const SubsurfaceExit& ex,
                                          RandomSampler& rand, const Lambdas& L, ShadowRay& sr) {
  sr.valid = false;
  uint32_t tri = ex.p.tri;
  float b1 = ex.p.b1, b2 = ex.p.b2, b0 = 1.f - b1 - b2;
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  vec3 p1 = load3(S.pos, i0), p2 = load3(S.pos, i1), p3 = load3(S.pos, i2);
  vec3 geometryNormal = normalize(cross(p2 - p1, p3 - p1));
  vec3 normal = normalize(b0 * load3(S.nrm, i0) + b1 * load3(S.nrm, i1) + b2 * load3(S.nrm, i2));
  vec3 tangent(1.f, 0.f, 0.f);
  if (S.tan) {
    const float* t0 = S.tan + 4 * i0;
    const float* t1 = S.tan + 4 * i1;
    const float* t2 = S.tan + 4 * i2;
    vec3 a(t0[0] * t0[3], t0[1] * t0[3], t0[2] * t0[3]);
    vec3 b(t1[0] * t1[3], t1[1] * t1[3], t1[2] * t1[3]);
    vec3 c(t2[0] * t2[3], t2[1] * t2[3], t2[2] * t2[3]);
    tangent = normalize(b0 * a + b1 * b + b2 * c);
  }
  vec3 wo = normalize(ex.p.n);
  SurfaceTexels tx;
  sample_surface_texels(S, m, tri, b1, b2, tx);
  SurfaceDef sd = make_surface(m, geometryNormal, normal, tangent, wo, true, (ps.flags & RS_REGULARIZE) != 0, L, S.spec, &tx);
  SurfaceContext ctx;
  make_surface_context(ctx, sd, current_ior(ps), previous_ior(ps), wo, ex.p.frontFace, L, false);
  if (hasLights) {
    float rl[4];
    rand.r4(rl);
    ExplicitLightSample ls;
    sample_explicit_light(S, fc, nee, ex.p.pos, rl, L, ls);
    if (!ls.emission.all_equal(0.f) && ls.pdf > 0.f) {
      Spectrum w;
      float bsdfPdf;
      evaluate_subsurface_exit_lobe(ctx, wo, normalize(ls.toLight), EVAL_NONE, w, bsdfPdf);
      if (!w.all_equal(0.f)) {
        sr.origin = spawn_ray_origin(ex.p.pos, ex.p.n, ls.toLight);
        sr.dir = ls.toLight;
        sr.tmax = ls.punctual ? ls.distance - RAY_MIN_T : ls.distance + RAY_MIN_T;
        sr.targetTri = ls.tri;
        float wMIS = ls.punctual ? 1.f : weight_mis(ls.pdf, bsdfPdf);
        w = (w / ls.pdf) * wMIS;
        sr.contrib = ps.throughput * w * ls.emission;
        sr.valid = true;
      }
    }
  }
  float re[4];
  rand.r4(re);
  vec3 wi = sample_subsurface_exit_lobe(ctx, wo, re[2], re[3]);
  Spectrum weight(0.f);
  float pdf = 0.f;
  vec3 nextDir(0.f), nextOrigin = ex.p.pos;
  if (!is_zero(wi)) {
    Spectrum w;
    evaluate_subsurface_exit_lobe(ctx, wo, wi, EVAL_NONE, w, pdf);
    if (pdf > 0.f) {
      weight = w / pdf;
      nextDir = normalize(wi);
      nextOrigin = spawn_ray_origin(ex.p.pos, ex.p.n, nextDir);
    }
  }
  ps.flags |= RS_REGULARIZE;
  ps.flags &= ~RS_DELTA;
  ps.pdfThisRay = pdf;
  ps.throughput = ps.throughput * weight;
  ps.dim = rand.dim;
  if (is_zero(nextDir) || pdf == 0.0f) {
    ps.rayState = RAY_CANCELLED;
  } else {
    ps.dir = nextDir;
    ps.origin = nextOrigin;
  }
}

// rayGenPrimaryRays.hlsl:72-101: after each traced segment, count it and apply Russian roulette.
YPT_HD void after_segment(const uint32_t* point, PathState& ps) {
  ps.pathLength += 1;
  if (ps.rayState != RAY_ALIVE) return;  // roulette on a finished path changes nothing it outputs
  RandomSampler r;
  r.point = point; r.dim = ps.dim; r.seed = ps.seed;
  if (ps.throughput.russian_roulette(r.r1())) { ps.rayState = RAY_TERMINATED; return; }
  ps.dim = r.dim;
}

// float4(totalLight.ToXYZ(), 1)
YPT_HD vec3 path_xyz(const SceneView& S, const PathState& ps) {
  return spectrum_to_xyz(ps.totalLight, spectral_lambdas(ps.lambda), S.spec);
}

}  // namespace ypt
