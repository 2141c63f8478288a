// Host preparation of scene and frame data (what the reference's managers upload each frame).
//
//  * materials: MaterialInternal::validateMaterial (MaterialInternal.cpp:27-43), the packing clamps of
//    BindlessMaterialManager.cpp:70-100, and the RGB -> spectrum coefficient conversion the hit
//    shaders perform per hit (convertSurfaceDefinitionFromRGB, hitShadersCommon.hlsl:51-91), done
//    once per material because untextured colours are constant;
//  * mesh lights: emissive-material discovery (LightManager.cpp:145), power estimates, selection
//    probabilities with a 10% uniform share and Vose alias tables (RaytraceCommonResources.cpp:325-547);
//  * frame constants: RaytraceCommonResources::update (RaytraceCommonResources.cpp:64-140).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/yapt_pt.h"
#include "../core/scene.h"
#include "rgb2spec.h"

namespace ypt {

struct PreparedScene {
  std::vector<float> pos, nrm, tan;  // tan: 4 per vertex (xyz * w applied at fetch), empty = no tangents
  std::vector<uint32_t> idx, triInst;
  std::vector<InstanceGPU> inst;
  std::vector<MaterialGPU> mats;
  std::vector<LightGPU> lights;
  uint32_t punctualCount = 0;  // the last entries of lights
  std::vector<float> uv;                // 2 per vertex, empty: no UVs
  std::vector<TextureGPU> textures;
  std::vector<uint32_t> texels;         // RGBA8
  std::vector<float> srgbLut;           // 256 entries, with textures
  std::vector<AlphaTri> alphaTris;      // alpha-tested triangles
  std::vector<uint32_t> triAlpha;       // per triangle: 0 = opaque, else 1 + its alphaTris record
  uint32_t envType = ENV_NONE, envWidth = 0, envHeight = 0;
  std::vector<float> envTexels;         // RGBA, one image or six cube faces
  float boundsMin[3], boundsMax[3];
};

inline float sat(float v) { return std::min(std::max(v, 0.f), 1.f); }

// setFromRGB (spectralDistribution.hlsl:183-197) as coefficients
inline void rgb_to_coeffs(Rgb2Spec& r2s, const float v[3], float c[3]) {
  if (v[0] == v[1] && v[0] == v[2]) { c[0] = 0.f; c[1] = 0.f; c[2] = sigmoid_inv(v[0]); return; }
  r2s.coeffs(v, c);
}
inline SpectralColor bounded_color(Rgb2Spec& r2s, const float v[3]) {
  SpectralColor s;
  rgb_to_coeffs(r2s, v, s.c);
  s.scale = 1.f;
  return s;
}
// setFromRGBUnbounded(NoIlluminant): coefficients of values / (2 max) and the scale 2 max
inline SpectralColor unbounded_color(Rgb2Spec& r2s, const float v[3]) {
  float m = std::max(std::max(v[0], v[1]), v[2]);
  float scale = 2.f * m;
  float n[3];
  for (int k = 0; k < 3; ++k) n[k] = scale == 0.f ? 0.f : v[k] / scale;
  SpectralColor s;
  rgb_to_coeffs(r2s, n, s.c);
  s.scale = scale;
  return s;
}

// the World.Skycube texture of a scene description, validated (RaytraceCommonResources.cpp:285-313)
inline void prepare_environment(const ypt_scene_desc& d, PreparedScene& out) {
  out.envType = ENV_NONE;
  out.envWidth = out.envHeight = 0;
  out.envTexels.clear();
  if (d.environment && d.environment->type != YPT_ENVIRONMENT_NONE) {
    const ypt_environment& e = *d.environment;
    if (e.type != YPT_ENVIRONMENT_CUBE && e.type != YPT_ENVIRONMENT_LONGLAT) throw std::invalid_argument("unknown environment type");
    if (e.width == 0 || e.height == 0 || !e.rgba) throw std::invalid_argument("environment without texels");
    if (e.type == YPT_ENVIRONMENT_CUBE && e.width != e.height) throw std::invalid_argument("cube map faces must be square");
    if (e.width > 65536u || e.height > 65536u) throw std::invalid_argument("environment larger than 65536 texels a side");
    size_t n = (size_t)e.width * e.height * 4u * (e.type == YPT_ENVIRONMENT_CUBE ? 6u : 1u);
    out.envTexels.assign(e.rgba, e.rgba + n);
    for (float v : out.envTexels)
      if (!(v >= 0.f) || !(v < INFINITY)) throw std::invalid_argument("environment texels must be finite and non-negative");
    out.envType = e.type == YPT_ENVIRONMENT_CUBE ? ENV_CUBE : ENV_LONGLAT;
    out.envWidth = e.width;
    out.envHeight = e.height;
  }
}

inline MaterialGPU convert_material(Rgb2Spec& r2s, const ypt_material& in, uint32_t textureCount) {
  ypt_material p = in;
  // validateMaterial
  p.roughness = std::max(0.001f, p.roughness);
  p.emission_focus = std::max(0.f, p.emission_focus);
  p.subsurface_amount = std::min(std::max(p.subsurface_amount, 0.f), 1.f);
  for (int k = 0; k < 3; ++k) p.subsurface_radius[k] = std::max(p.subsurface_radius[k], 0.f);
  p.subsurface_radius_scale = std::max(0.f, p.subsurface_radius_scale);
  float coatingSum = p.clear_coat_amount + p.sheen_amount;
  if (coatingSum > 1) {
    float inv = 1.f / coatingSum;
    p.clear_coat_amount *= inv;
    p.sheen_amount *= inv;
  }
  MaterialGPU m;
  std::memset(&m, 0, sizeof(m));
  m.specularAmount = sat(p.specular_amount);
  m.clearCoatAmount = sat(p.clear_coat_amount);
  m.clearCoatIOR = p.clear_coat_ior;
  m.clearCoatRoughness = sat(p.clear_coat_roughness);
  m.transparency = p.transparency;
  m.metalness = p.metalness;
  m.dielectricIOR = std::min(std::max(p.dielectric_ior, 0.3f), 3.0f);
  m.roughness = sat(p.roughness);
  m.anisotropy = sat(p.anisotropy);
  m.anisotropyRotation = p.anisotropy_rotation;
  m.flags = p.material_mask;
  m.thinFilmThicknessNM = p.thin_film_thickness_nm;
  m.cauchyA = p.cauchy_coeffs[0];
  m.cauchyB = p.cauchy_coeffs[1];
  // extensions: a blackbody below 500 K emits nothing visible in 380-700 nm at float precision
  m.blackbodyK = p.blackbody_kelvin > 0.f ? std::max(p.blackbody_kelvin, 500.f) : 0.f;
  for (int k = 0; k < 3; ++k) {
    m.sellmeierB[k] = p.sellmeier_b[0] > 0.f ? p.sellmeier_b[k] : 0.f;
    m.sellmeierC[k] = p.sellmeier_b[0] > 0.f ? p.sellmeier_c[k] : 0.f;
  }
  m.sheenAmount = p.sheen_amount;
  m.emissionFocus = p.emission_focus;
  m.alphaCutoff = p.alpha_cutoff;
  m.sheenRoughness = p.sheen_roughness;
  m.subsurfaceAmount = sat(p.subsurface_amount);
  // texture bindings: indices outside [0, texture count) are unbound (TEX_UNBOUND_INDEX); the UV scale
  // is kept at half precision (unpackTextureTransformScale)
  const int32_t* texIdx[4] = {&p.albedo_tex, &p.normal_tex, &p.orm_tex, &p.emissive_tex};
  const float* texScale[4] = {p.albedo_tex_scale, p.normal_tex_scale, p.orm_tex_scale, p.emissive_tex_scale};
  for (int k = 0; k < 4; ++k) {
    m.tex[k] = (*texIdx[k] >= 0 && (uint32_t)*texIdx[k] < textureCount) ? *texIdx[k] : -1;
    m.texScale[2 * k] = f16_to_f32(f32_to_f16(texScale[k][0]));
    m.texScale[2 * k + 1] = f16_to_f32(f32_to_f16(texScale[k][1]));
  }
  float ssr[3];
  for (int k = 0; k < 3; ++k) ssr[k] = p.subsurface_radius[k] * p.subsurface_radius_scale;
  m.specular = bounded_color(r2s, p.specular);
  m.albedo = bounded_color(r2s, p.albedo);
  m.sheenColor = bounded_color(r2s, p.sheen_tint);
  m.absorption = unbounded_color(r2s, p.absorption);
  m.emissive = unbounded_color(r2s, p.emissive);
  m.subsurfaceRadius = unbounded_color(r2s, ssr);
  m.emissiveNonZero = (p.emissive[0] + p.emissive[1] + p.emissive[2]) > 0.f ? 1u : 0u;
  for (int k = 0; k < 3; ++k) {
    m.albedoRGB[k] = p.albedo[k];
    m.emissiveRGB[k] = p.emissive[k];
    m.sheenRGB[k] = p.sheen_tint[k];
  }
  return m;
}

inline void selection_pdfs(std::vector<float>& w) {  // calculateSelectionPDFs
  const float UNIFORM_SHARE = 0.1f;
  size_t count = w.size();
  double total = 0;
  for (float x : w) total += x;
  float u = total > 0 ? UNIFORM_SHARE : 1.f;
  for (float& x : w) {
    float share = total > 0 ? float(x / total) : 0.f;
    x = (1.f - u) * share + u / count;
  }
}
inline void alias_table(const std::vector<float>& pdfs, std::vector<float>& thr, std::vector<uint32_t>& ali) {  // buildAliasTable
  size_t count = pdfs.size();
  thr.resize(count);
  ali.resize(count);
  std::vector<float> scaled(count);
  std::vector<uint32_t> small, large;
  for (size_t i = 0; i < count; ++i) {
    scaled[i] = pdfs[i] * count;
    (scaled[i] < 1.f ? small : large).push_back(uint32_t(i));
  }
  while (!small.empty() && !large.empty()) {
    uint32_t s = small.back(); small.pop_back();
    uint32_t l = large.back(); large.pop_back();
    thr[s] = scaled[s];
    ali[s] = l;
    scaled[l] = (scaled[l] + scaled[s]) - 1.f;
    (scaled[l] < 1.f ? small : large).push_back(l);
  }
  for (uint32_t i : large) { thr[i] = 1.f; ali[i] = i; }
  for (uint32_t i : small) { thr[i] = 1.f; ali[i] = i; }
}

inline void prepare_scene(const ypt_scene_desc& d, Rgb2Spec& r2s, PreparedScene& out) {
  if (!d.positions || !d.indices || d.triangle_count == 0 || d.vertex_count == 0)
    throw std::invalid_argument("scene has no geometry");
  if (!d.materials || d.material_count == 0) throw std::invalid_argument("scene has no materials");
  if (d.instance_count == 0 || !d.instance_first_triangle || !d.instance_triangle_count || !d.instance_material)
    throw std::invalid_argument("scene has no instances");
  out.pos.assign(d.positions, d.positions + 3 * (size_t)d.vertex_count);
  out.idx.assign(d.indices, d.indices + 3 * (size_t)d.triangle_count);
  for (size_t k = 0; k < out.idx.size(); ++k)
    if (out.idx[k] >= d.vertex_count) throw std::invalid_argument("vertex index out of range");
  out.nrm.resize(out.pos.size());
  if (d.normals) {
    std::memcpy(out.nrm.data(), d.normals, out.nrm.size() * sizeof(float));
  } else {  // per-vertex copy of the face normal requires unshared vertices; use the area-weighted average
    std::fill(out.nrm.begin(), out.nrm.end(), 0.f);
    for (uint32_t t = 0; t < d.triangle_count; ++t) {
      const float* a = &out.pos[3 * out.idx[3 * t]];
      const float* b = &out.pos[3 * out.idx[3 * t + 1]];
      const float* c = &out.pos[3 * out.idx[3 * t + 2]];
      float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      for (int k = 0; k < 3; ++k)
        for (int a2 = 0; a2 < 3; ++a2) out.nrm[3 * out.idx[3 * t + k] + a2] += n[a2];
    }
    for (uint32_t v = 0; v < d.vertex_count; ++v) {
      float* n = &out.nrm[3 * v];
      float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (l > 0.f) { n[0] /= l; n[1] /= l; n[2] /= l; } else { n[1] = 1.f; }
    }
  }
  out.tan.clear();
  if (d.tangents) out.tan.assign(d.tangents, d.tangents + 4 * (size_t)d.vertex_count);
  out.mats.resize(d.material_count);
  {  // colour fits are independent per material; the node cache is shared and locked
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        for (uint32_t m = t; m < d.material_count; m += nt) out.mats[m] = convert_material(r2s, d.materials[m], d.texture_count);
      });
    for (auto& x : th) x.join();
  }

  out.uv.clear();
  if (d.uvs) out.uv.assign(d.uvs, d.uvs + 2 * (size_t)d.vertex_count);
  out.textures.clear();
  out.texels.clear();
  out.srgbLut.clear();
  if (d.texture_count && !d.textures) throw std::invalid_argument("textures is null");
  for (uint32_t t = 0; t < d.texture_count; ++t) {
    const ypt_texture& tx = d.textures[t];
    if (tx.width == 0 || tx.height == 0 || !tx.rgba8) throw std::invalid_argument("texture without texels");
    if (out.texels.size() + (size_t)tx.width * tx.height > 0xFFFFFFFFull) throw std::invalid_argument("textures exceed 4G texels");
    out.textures.push_back(TextureGPU{(uint32_t)out.texels.size(), tx.width, tx.height, tx.srgb ? 1u : 0u});
    size_t n = (size_t)tx.width * tx.height;
    size_t base = out.texels.size();
    out.texels.resize(base + n);
    std::memcpy(out.texels.data() + base, tx.rgba8, n * 4);
  }
  prepare_environment(d, out);
  if (d.texture_count) {  // IEC 61966-2-1 decoding of each byte value
    out.srgbLut.resize(256);
    for (int b = 0; b < 256; ++b) {
      double c = b / 255.0;
      out.srgbLut[b] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
  }

  out.triInst.assign(d.triangle_count, 0xFFFFFFFFu);
  out.inst.resize(d.instance_count);
  for (uint32_t i = 0; i < d.instance_count; ++i) {
    uint32_t f = d.instance_first_triangle[i], c = d.instance_triangle_count[i];
    if (c == 0 || (uint64_t)f + c > d.triangle_count) throw std::invalid_argument("instance triangle range out of bounds");
    if (d.instance_material[i] >= d.material_count) throw std::invalid_argument("instance material out of range");
    for (uint32_t t = f; t < f + c; ++t) {
      if (out.triInst[t] != 0xFFFFFFFFu) throw std::invalid_argument("triangle claimed by two instances");
      out.triInst[t] = i;
    }
    out.inst[i] = InstanceGPU{d.instance_material[i], f, c, 0.f};
  }
  for (uint32_t t = 0; t < d.triangle_count; ++t)
    if (out.triInst[t] == 0xFFFFFFFFu) throw std::invalid_argument("triangle without instance");

  // alpha testing needs a cutoff, an albedo texture and UVs (hitShadersCommon.hlsl:166-175)
  out.alphaTris.clear();
  out.triAlpha.assign(d.triangle_count, 0u);
  if (!out.uv.empty()) {
    for (uint32_t t = 0; t < d.triangle_count; ++t) {
      const MaterialGPU& m = out.mats[out.inst[out.triInst[t]].material];
      if (!(m.alphaCutoff > 0.f) || m.tex[TEX_ALBEDO] < 0) continue;
      AlphaTri a{};
      for (int k = 0; k < 3; ++k) {
        a.uv[2 * k] = out.uv[2 * out.idx[3 * t + k]];
        a.uv[2 * k + 1] = out.uv[2 * out.idx[3 * t + k] + 1];
      }
      a.scale[0] = m.texScale[2 * TEX_ALBEDO];
      a.scale[1] = m.texScale[2 * TEX_ALBEDO + 1];
      a.tex = (uint32_t)m.tex[TEX_ALBEDO];
      a.cutoff = m.alphaCutoff;
      out.alphaTris.push_back(a);
      out.triAlpha[t] = (uint32_t)out.alphaTris.size();
    }
  }

  for (int a = 0; a < 3; ++a) { out.boundsMin[a] = 3.4e38f; out.boundsMax[a] = -3.4e38f; }
  for (size_t v = 0; v < d.vertex_count; ++v)
    for (int a = 0; a < 3; ++a) {
      out.boundsMin[a] = std::min(out.boundsMin[a], out.pos[3 * v + a]);
      out.boundsMax[a] = std::max(out.boundsMax[a], out.pos[3 * v + a]);
    }

  // mesh lights
  out.lights.clear();
  std::vector<float> power;
  for (uint32_t i = 0; i < d.instance_count; ++i) {
    const ypt_material& mp = d.materials[d.instance_material[i]];
    if (!(mp.emissive[0] + mp.emissive[1] + mp.emissive[2] > 0)) continue;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t t = out.inst[i].firstTri; t < out.inst[i].firstTri + out.inst[i].triCount; ++t)
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], out.pos[3 * out.idx[3 * t + k] + a]);
          hi[a] = std::max(hi[a], out.pos[3 * out.idx[3 * t + k] + a]);
        }
    float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float area = ex * ey + ey * ez + ez * ex;
    float emission = 0.2126f * mp.emissive[0] + 0.7152f * mp.emissive[1] + 0.0722f * mp.emissive[2];
    if (emission <= 0.f) emission = 1.f;
    float sides = (mp.material_mask & YPT_MATERIAL_TWO_SIDED) ? 2.f : 1.f;
    power.push_back(3.14159265358979323846f * emission * area * sides);
    LightGPU L;
    std::memset(&L, 0, sizeof(L));
    L.instance = i;
    L.material = d.instance_material[i];
    L.firstTri = out.inst[i].firstTri;
    L.triCount = out.inst[i].triCount;
    L.type = LIGHT_TYPE_MESH;
    out.lights.push_back(L);
  }
  size_t meshLights = out.lights.size();
  out.punctualCount = d.punctual_light_count;
  if (d.punctual_light_count && !d.punctual_lights) throw std::invalid_argument("punctual_lights is null");
  std::vector<float> punctualPower;
  for (uint32_t k = 0; k < d.punctual_light_count; ++k) {  // uploadLightData (RaytraceCommonResources.cpp:518-541)
    const ypt_punctual_light& p = d.punctual_lights[k];
    if (p.type < YPT_LIGHT_POINT || p.type > YPT_LIGHT_DIRECTIONAL) throw std::invalid_argument("unknown punctual light type");
    LightGPU L;
    std::memset(&L, 0, sizeof(L));
    L.instance = 0xFFFFFFFFu;
    L.type = p.type;
    float dl = std::sqrt(p.direction[0] * p.direction[0] + p.direction[1] * p.direction[1] + p.direction[2] * p.direction[2]);
    float dir[3] = {0.f, 0.f, -1.f};
    if (dl > 0.f)
      for (int a = 0; a < 3; ++a) dir[a] = p.direction[a] / dl;
    for (int a = 0; a < 3; ++a) {
      L.positionCosInner[a] = p.position[a];
      L.directionCosOuter[a] = dir[a];
      L.intensity[a] = std::max(0.f, p.intensity[a]);
    }
    L.positionCosInner[3] = p.cos_inner;
    L.directionCosOuter[3] = p.cos_outer;
    // estimatePunctualLightPower (RaytraceCommonResources.cpp:407-423)
    float lum = 0.2126f * L.intensity[0] + 0.7152f * L.intensity[1] + 0.0722f * L.intensity[2];
    float pw = lum;
    if (p.type == YPT_LIGHT_POINT) pw = 4.f * 3.14159265358979323846f * lum;
    else if (p.type == YPT_LIGHT_SPOT) pw = 2.f * 3.14159265358979323846f * (1.f - 0.5f * (p.cos_inner + p.cos_outer)) * lum;
    else {
      float ex = (out.boundsMax[0] - out.boundsMin[0]) * 0.5f, ey = (out.boundsMax[1] - out.boundsMin[1]) * 0.5f,
            ez = (out.boundsMax[2] - out.boundsMin[2]) * 0.5f;
      pw = 3.14159265358979323846f * (ex * ex + ey * ey + ez * ez) * lum;
    }
    power.push_back(pw);
    punctualPower.push_back(pw);
    out.lights.push_back(L);
  }
  if (!out.lights.empty()) {
    std::vector<float> pdfs = power, thr;
    std::vector<uint32_t> ali;
    selection_pdfs(pdfs);
    alias_table(pdfs, thr, ali);
    for (size_t l = 0; l < out.lights.size(); ++l) {
      out.lights[l].selectionPdf = pdfs[l];
      out.lights[l].aliasThreshold = thr[l];
      out.lights[l].aliasIndex = ali[l];
      if (l < meshLights) out.inst[out.lights[l].instance].lightSelectionPdf = pdfs[l];
    }
  }
  if (!punctualPower.empty()) {
    std::vector<float> pdfs = punctualPower, thr;
    std::vector<uint32_t> ali;
    selection_pdfs(pdfs);
    alias_table(pdfs, thr, ali);
    for (size_t k = 0; k < pdfs.size(); ++k) {
      LightGPU& L = out.lights[meshLights + k];
      L.punctualSelectionPdf = pdfs[k];
      L.punctualAliasThreshold = thr[k];
      L.punctualAliasIndex = ali[k];
    }
  }
}

inline void prepare_frame(const ypt_frame_desc& f, const PreparedScene& s, uint32_t integrator, FrameConstants& fc) {
  if (f.width == 0 || f.height == 0) throw std::invalid_argument("frame has zero size");
  if (f.max_depth == 0) throw std::invalid_argument("max_depth must be positive");
  std::memset(&fc, 0, sizeof(fc));
  std::memcpy(fc.uvToView.m, f.uv_to_view, 64);
  std::memcpy(fc.viewToUV.m, f.view_to_uv, 64);
  std::memcpy(fc.worldToView.m, f.world_to_view, 64);
  std::memcpy(fc.viewToWorld.m, f.view_to_world, 64);
  for (int a = 0; a < 3; ++a) { fc.worldBoundsMin[a] = s.boundsMin[a]; fc.worldBoundsMax[a] = s.boundsMax[a]; }
  fc.cameraPosition[0] = f.view_to_world[3];
  fc.cameraPosition[1] = f.view_to_world[7];
  fc.cameraPosition[2] = f.view_to_world[11];
  fc.cameraPosition[3] = 1.f;
  float pw = 1.f / f.width, ph = 1.f / f.height;
  fc.targetTexDimensions[0] = (float)f.width;
  fc.targetTexDimensions[1] = (float)f.height;
  fc.targetTexDimensions[2] = pw;
  fc.targetTexDimensions[3] = ph;
  fc.rayUVOffset[0] = (float)(0.5 * pw);
  fc.rayUVOffset[1] = (float)(0.5 * ph);
  fc.maxRayDepth = f.max_depth;
  fc.flags = f.flags;
  fc.envIntensityScale = f.env_intensity_scale;
  fc.lightCount = (uint32_t)s.lights.size();
  fc.punctualLightCount = s.punctualCount;
  fc.regularizeAfterVertices = f.regularize_after;
  fc.integrator = integrator;
  fc.lensParams[0] = std::max(0.f, f.aperture_radius);
  fc.lensParams[1] = std::max(1e-4f, f.focus_distance);
  fc.lensParams[2] = f.forward_z_sign;
  fc.lensParams[3] = 0.f;
  uint32_t lv = f.light_path_vertices == 0 ? 16u : f.light_path_vertices;
  if (lv < 2 || lv > 16) throw std::invalid_argument("light_path_vertices must be in [2, 16]");
  fc.lightPathVertices = lv;
  fc.lightPathsPerDim[0] = (f.width + 7u) / 8u;   // one light subpath per 8x8 texels (TEXELS_PER_LIGHTPATH)
  fc.lightPathsPerDim[1] = (f.height + 7u) / 8u;
  // _DEBUG.BDPT.EffectiveLightPathNodes / EffectiveCameraPathNodes, default (0, 128) each (RendererVarsList.h:46-47)
  const bool ranges = (f.flags & YPT_FLAG_BDPT_NODE_RANGES) != 0;
  for (int k = 0; k < 4; ++k) fc.nodeRanges[k] = ranges ? f.bdpt_node_ranges[k] : (k & 1 ? 128u : 0u);
}

}  // namespace ypt
