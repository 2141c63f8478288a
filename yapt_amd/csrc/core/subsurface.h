// Subsurface scattering: the Christensen-Burley diffusion profile as a separable BSSRDF,
// S(po, wo, pi, wi) = entry(po, wo) * Rd(|po - pi|) * Sw(pi, wi) (subsurface.hlsl:1-392,
// scatteringSample.hlsl:30-78). A path that takes the subsurface layer enters at po and continues
// from an exit point pi found by probe rays against the same instance (King et al. 2013): pick a
// probe axis (normal 1/2, tangent 1/4, bitangent 1/4) and a spectral lane, sample a radius from
// that lane's profile, and project the disk point along the axis onto the instance, choosing among
// up to 8 hits in proportion to the profile. The exit leaves through a cosine lobe.
//
// The probe needs a closest-hit query restricted to a triangle range (an instance's triangles);
// it comes in as a Tracer with
//   bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h)
// so the GPU kernels and the CPU checker run the same code.
#pragma once
#include "material.h"
#include "scene.h"

namespace ypt {

static constexpr int SSS_LANES = 4;
static constexpr int SSS_MAX_PROBE_HITS = 8;
static constexpr float SSS_MIN_D = 1e-6f;
static constexpr float SSS_RMAX_FACTOR = 20.7232658f;  // 3 ln(1000): the 0.999 tail of the slow component

// getSubsurfaceLayerWeight (subsurface.hlsl:25-28)
YPT_HD float subsurface_layer_weight(const SurfaceDef& s) {
  return (1.f - s.metalness) * (1.f - s.transparency) * s.subsurfaceAmount;
}

// getEnergyLeftForBaseLayer (subsurface.hlsl:31-73): what the coating, sheen, conductor and
// specular layers above the base leave, for the outgoing direction
YPT_HD float energy_left_for_base_layer(const SurfaceContext& c, const MSLuts& Lu) {
  const SurfaceDef& s = c.def;
  const PrecalcSurface& p = c.pre;
  float energyLeft = 1.f;
  float fromIOR = p.fromIOR, toIOR = p.toIOR;
  if (s.clearCoatAmount > 0.f)
    energyLeft *= energy_remaining_after_specular(Lu, s.clearCoatIOR / fromIOR, p.woCoating.y, s.clearCoatRoughness,
                                                  s.clearCoatAmount);
  if (s.sheenAmount > 0.f) energyLeft *= energy_remaining_after_sheen(Lu, p.woBase.y, s.sheenRoughness, s.sheenAmount);
  if (s.metalness > 0.f) energyLeft *= 1.f - s.metalness;
  float specAmount = (1.f - s.metalness) * s.specularAmount;
  if (specAmount > 0.f) {
    if (has_dispersion(s.flags)) {
      if (p.exiting) fromIOR = dispersed_ior(s, c.lambdas);
      else toIOR = dispersed_ior(s, c.lambdas);
    }
    SpecularLayer sl = make_specular_layer(p.ax, p.ay, s.roughness, toIOR / fromIOR, true, s.transparency > 0.f);
    energyLeft *= sl.energy_left(Lu, p.woBase, p.woBase, specAmount);
  }
  return energyLeft;
}

// getSubsurfaceEntryFactor (subsurface.hlsl:74-82)
YPT_HD float subsurface_entry_factor(const SurfaceContext& c, const MSLuts& Lu) {
  return subsurface_layer_weight(c.def) * energy_left_for_base_layer(c, Lu) * c.def.diffOcclusion;
}

// Burley profile, normalised over the plane: R(r) = (e^(-r/d) + e^(-r/(3d))) / (8 pi d r), with the
// radius taken as the diffuse mean free path and d = radius / s(albedo) (subsurface.hlsl:85-143)
YPT_HD float burley_scale(float albedo) {
  float x = fabsf(albedo - 0.8f);
  return 1.85f - albedo + 7.f * x * x * x;
}
YPT_HD float burley_d(float radius, float albedo) { return fmax_(radius, SSS_MIN_D) / burley_scale(albedo); }
YPT_HD float burley_rmax(float d) { return SSS_RMAX_FACTOR * d; }
YPT_HD float burley_profile_area(float r, float d) {
  r = fmax_(r, d * 1e-4f);
  return (yexp(-r / d) + yexp(-r / (3.f * d))) / (8.f * PI * d * r);
}
// the radial density is the mixture (1/4) e^(-r/d)/d + (3/4) e^(-r/(3d))/(3d)
YPT_HD float burley_sample_radius(float u, float d) {
  float scale = d;
  if (u < 0.25f) {
    u *= 4.f;
  } else {
    u = (u - 0.25f) / 0.75f;
    scale = 3.f * d;
  }
  return -scale * ylog(fmax_(1.f - u, 1e-7f));
}

// Rd(distance) for the four lanes (evaluateSubsurfaceRd)
YPT_HD Spectrum subsurface_rd(const Spectrum& albedo, const Spectrum& radius, float distance) {
  Spectrum rd;
  for (int k = 0; k < SSS_LANES; ++k)
    rd.s[k] = albedo.s[k] * burley_profile_area(distance, burley_d(radius.s[k], albedo.s[k]));
  return rd;
}

// getSubsurfaceProbeFrame: axis 0 = normal, 1 = tangent, 2 = bitangent of the normal's own basis, so
// the pdf is a function of positions and normals alone
YPT_HD void subsurface_probe_frame(vec3 n, int axisIndex, vec3& axis, vec3& planeU, vec3& planeV) {
  vec3 t, b;
  construct_vector_base(n, t, b);
  if (axisIndex == 0) { axis = n; planeU = t; planeV = b; }
  else if (axisIndex == 1) { axis = t; planeU = b; planeV = n; }
  else { axis = b; planeU = n; planeV = t; }
}
YPT_HD float subsurface_axis_probability(int axisIndex) { return axisIndex == 0 ? 0.5f : 0.25f; }

// pdfSubsurfaceExit (subsurface.hlsl:177-200): area density of `to` from `from`, one-sample MIS over
// the axes and lanes; the choice among probe hits is not part of it
YPT_HD float pdf_subsurface_exit(const Spectrum& albedo, const Spectrum& radius, vec3 fromPos, vec3 fromN, vec3 toPos,
                                 vec3 toN) {
  vec3 d = toPos - fromPos;
  float pdf = 0.f;
  for (int a = 0; a < 3; ++a) {
    vec3 axis, pu, pv;
    subsurface_probe_frame(fromN, a, axis, pu, pv);
    float du = dot(d, pu), dv = dot(d, pv);
    float rProj = sqrtf(du * du + dv * dv);
    float axisPdf = 0.f;
    for (int k = 0; k < SSS_LANES; ++k) {
      float dl = burley_d(radius.s[k], albedo.s[k]);
      if (rProj < burley_rmax(dl)) axisPdf += burley_profile_area(rProj, dl);
    }
    pdf += subsurface_axis_probability(a) * (1.f / SSS_LANES) * axisPdf * fabsf(dot(toN, axis));
  }
  return pdf;
}

struct SurfacePoint {
  vec3 pos;
  vec3 n;           // geometry normal on the side the path enters from or leaves to
  uint32_t inst;
  uint32_t tri;
  float b1, b2;
  bool frontFace;
};

struct SubsurfaceExit {
  SurfacePoint p;
  Spectrum weight;    // Rd / (pdfPosition * probe-hit selection probability)
  float pdfPosition;  // area density of the exit point (pdf_subsurface_exit)
};

// sampleSubsurfaceExit (subsurface.hlsl:203-339)
template <class Tracer>
YPT_HD bool sample_subsurface_exit(const SceneView& S, const Spectrum& albedo, const Spectrum& radius,
                                   const SurfacePoint& entry, float r0, float r1, float r2, Tracer& tr,
                                   SubsurfaceExit& out) {
  float u = r0;
  int axisIndex;
  if (u < 0.5f) { axisIndex = 0; u *= 2.f; }
  else if (u < 0.75f) { axisIndex = 1; u = (u - 0.5f) * 4.f; }
  else { axisIndex = 2; u = (u - 0.75f) * 4.f; }
  uint32_t lane = (uint32_t)(u * (float)SSS_LANES);
  if (lane > SSS_LANES - 1) lane = SSS_LANES - 1;
  float uSelect = saturate(u * (float)SSS_LANES - (float)lane);
  float dl = burley_d(sel_get(radius.s, lane), sel_get(albedo.s, lane));
  float rMax = burley_rmax(dl);
  float r = burley_sample_radius(r1, dl);
  if (r >= rMax) return false;
  vec3 axis, pu, pv;
  subsurface_probe_frame(entry.n, axisIndex, axis, pu, pv);
  float sinPhi, cosPhi;
  ysincos(2.f * PI * r2, &sinPhi, &cosPhi);
  // chord of the r_max sphere through the disk point, kept above a minimum length so that a tiny
  // profile does not lose its hits to float precision
  vec3 e = entry.pos;
  float maxCoord = fmax_(fmax_(fabsf(e.x), fabsf(e.y)), fabsf(e.z));
  float halfLength = fmax_(sqrtf(rMax * rMax - r * r), 1e-3f + 1e-5f * maxCoord);
  vec3 probeOrigin = e + r * (cosPhi * pu + sinPhi * pv) - axis * halfLength;
  const InstanceGPU& inst = S.inst[entry.inst];
  float wSum = 0.f, wSel = 0.f, selT = 0.f, selB1 = 0.f, selB2 = 0.f;
  Spectrum selRd(0.f);
  uint32_t selTri = 0;
  float tMin = 0.f;
  for (int i = 0; i < SSS_MAX_PROBE_HITS; ++i) {
    Hit h;
    if (!tr.closest_in_range(probeOrigin, axis, tMin, 2.f * halfLength, inst.firstTri, inst.triCount, h)) break;
    // keep the k-th hit with probability w_k / (w_1 + ... + w_k), without storing the hits
    Spectrum rd = subsurface_rd(albedo, radius, length(probeOrigin + axis * h.t - e));
    float w = rd.s[0] + rd.s[1] + rd.s[2] + rd.s[3];
    wSum += w;
    float keep = wSum > 0.f ? w / wSum : 0.f;
    if (keep >= 1.f || uSelect < keep) {
      uSelect = keep >= 1.f ? uSelect : uSelect / keep;
      wSel = w; selRd = rd; selT = h.t; selTri = h.tri; selB1 = h.b1; selB2 = h.b2;
    } else {
      uSelect = (uSelect - keep) / (1.f - keep);
    }
    tMin = h.t + RAY_MIN_T;
  }
  if (!(wSel > 0.f)) return false;
  vec3 p1, p2, p3;
  tri_vertices(S, selTri, p1, p2, p3);
  vec3 c = cross(p2 - p1, p3 - p1);
  vec3 nWS = normalize(c);
  bool selFront = dot(axis, c) < 0.f;  // the probe arrives on the triangle's front side
  // orient by the side the probe hit, independent of the winding
  vec3 frontN = ((dot(nWS, axis) < 0.f) == selFront) ? nWS : -nWS;
  out.p.pos = probeOrigin + axis * selT;
  out.p.inst = entry.inst;
  out.p.n = entry.frontFace ? frontN : -frontN;  // leave to the side the path entered from
  out.p.tri = selTri;
  out.p.b1 = selB1;
  out.p.b2 = selB2;
  out.p.frontFace = entry.frontFace;
  float pdf = pdf_subsurface_exit(albedo, radius, e, entry.n, out.p.pos, out.p.n);
  if (!(pdf > 0.f)) return false;
  float selProb = wSel / wSum;
  out.weight = selRd / (pdf * selProb);
  out.pdfPosition = pdf;
  return true;
}

// sampleSubsurfaceExitLobe / evaluateSubsurfaceExitLobe (subsurface.hlsl:342-392): cosine lobe about
// the base normal, never to the inside of the geometry
YPT_HD vec3 sample_subsurface_exit_lobe(const SurfaceContext& c, vec3 wo, float r0, float r1) {
  vec3 wi = mul(sample_cosine_hemisphere(r0, r1), c.pre.to_base(c.def.tangent));
  if (dot(wi, c.def.geometryNormal) * dot(wo, c.def.geometryNormal) <= 0.f) return vec3(0.f);
  return wi;
}
YPT_HD void evaluate_subsurface_exit_lobe(const SurfaceContext& c, vec3 wo, vec3 wi, uint32_t flags, Spectrum& weight,
                                          float& pdf) {
  weight.set(0.f);
  pdf = 0.f;
  float dWoG = dot(wo, c.def.geometryNormal), dWiG = dot(wi, c.def.geometryNormal);
  if (dWoG * dWiG <= 0.f) return;
  vec3 wiBase = normalize(mul(c.pre.to_base(c.def.tangent), wi));
  if (wiBase.y <= 0.f) return;
  pdf = pdf_cosine_hemisphere(wiBase.y);
  if ((flags & EVAL_PDF_ONLY) == 0) {
    // Sw does not depend on the arriving direction: on a light path the shading cosine becomes
    // the geometric one of the leaving direction
    float cosine = (flags & EVAL_LIGHT_PATH) != 0 ? fabsf(dWiG) : wiBase.y;
    weight.set(INVPI * cosine);
  }
}

}  // namespace ypt
