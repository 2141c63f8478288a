// Bidirectional path tracing with all-strategy MIS: light subpaths, camera subpaths connected to
// every vertex of one light subpath, and light tracing onto the lens (s >= 2, t = 1).
//
// Restates shaders/raytracing/BDPTCommon.hlsl:26-1223 (node layout, BDPTRayState,
// evaluateSurfaceAndPDFs, calculateCameraDirectionalPDF, calculatePDFRatio, accumulateRISum,
// calculateConnectingLightNodeWeightsAndPDFs, calculateMISWeight*, calculateLightPathRadiance),
// BDPTTraceLightPath.comp:6-362, BDPTTraceCameraPath.comp:40-432 and
// BDPTCalculatePureLightPaths.comp:8-206 for mesh lights. Energy compensation is the reciprocal
// (Kulla-Conty) form (RECIPROCAL_MULTISCATTER).
//
// The tracing functions are templates over a Tracer with
//   bool closest(vec3 o, vec3 d, float tmin, float tmax, Hit& h)         closest (t, id) in [tmin, tmax]
//   bool hits_target(vec3 o, vec3 d, float tmin, float tmax, uint32_t tri) the closest hit is `tri`
//   bool unoccluded(vec3 o, vec3 d, float tmin, float tmax)              nothing is hit in [tmin, tmax]
//   bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h)
//                                                                        subsurface probes (one instance)
// so that the GPU kernels and the CPU checker run the same vertex arithmetic over their own BVHs.
//
// Subsurface scattering: a subpath that takes the subsurface event continues from the exit point as
// its next vertex, with no ray in between; the edge's densities are area densities, and exit
// vertices cannot be reached by a connection, so they stay out of the MIS sums where the reference
// keeps them out (accumulate_ri's `connectable`).
#pragma once
#include "shade.h"

namespace ypt {

static constexpr uint32_t BDPT_MAX_LIGHT_VERTICES = 16;  // MAX_LIGHT_PATH_VERTICES_HARD_LIMIT
static constexpr uint32_t BDPT_TILE = 8;                  // THREAD_GROUP_DIM_SIZE_CAM_RAY = TEXELS_PER_LIGHTPATH
static constexpr float BDPT_MAX_PDF_RATIO_SUM = 1e16f;
static constexpr float BDPT_VISIBILITY_EPS = 0.001f;

// LIGHT_PATH_NODE_FLAG_* (BDPTCommon.hlsl:14-28)
static constexpr uint32_t LN_SECONDARY_TERMINATED = 1u << 0;
static constexpr uint32_t LN_FRONT_FACE = 1u << 1;
static constexpr uint32_t LN_ENV = 1u << 2;
static constexpr uint32_t LN_DELTA = 1u << 3;
static constexpr uint32_t LN_REGULARIZED = 1u << 4;
static constexpr uint32_t LN_SUBSURFACE_EXIT = 1u << 5;  // reached through the subsurface: no ray before it
static constexpr uint32_t LN_PUNCTUAL = 1u << 6;
static constexpr uint32_t LN_DIRECTIONAL = 1u << 7;

// ---------------------------------------------------------------- light subpath nodes
// LightPathNodePacked (BDPTCommon.hlsl:37-44, 300-335): 64 bytes here (no next-index link: the
// vertices of a subpath are stored in order, light first).
struct LightNodePacked {
  float radiance[4];
  float pos[3];
  uint32_t tri;        // global triangle id (instance + primitive)
  float b1, b2;        // barycentrics; (-1, -1) for the light origin
  float pdfFwd, pdfBwd;
  uint32_t nxy;        // f16 normal x | y << 16
  uint32_t nzFlags;    // f16 normal z | flags << 16
  float iorPrev, iorCur;
};

struct LightNode {     // LightPathNode (:46-60)
  Spectrum radiance;
  vec3 pos, nrm;
  uint32_t tri;
  float b1, b2, pdfFwd, pdfBwd, iorPrev, iorCur;
  uint32_t flags;
};

YPT_HD LightNodePacked pack_node(const LightNode& n) {
  LightNodePacked p;
  for (int i = 0; i < 4; ++i) p.radiance[i] = n.radiance.s[i];
  p.pos[0] = n.pos.x; p.pos[1] = n.pos.y; p.pos[2] = n.pos.z;
  p.tri = n.tri;
  p.b1 = n.b1; p.b2 = n.b2;
  p.pdfFwd = n.pdfFwd; p.pdfBwd = n.pdfBwd;
  p.nxy = f32_to_f16(n.nrm.x) | (f32_to_f16(n.nrm.y) << 16);
  p.nzFlags = f32_to_f16(n.nrm.z) | (n.flags << 16);
  p.iorPrev = n.iorPrev; p.iorCur = n.iorCur;
  return p;
}
YPT_HD LightNode unpack_node(const LightNodePacked& p) {
  LightNode n;
  n.radiance = Spectrum(p.radiance[0], p.radiance[1], p.radiance[2], p.radiance[3]);
  n.pos = vec3(p.pos[0], p.pos[1], p.pos[2]);
  n.nrm = vec3(f16_to_f32(p.nxy & 0xFFFFu), f16_to_f32(p.nxy >> 16), f16_to_f32(p.nzFlags & 0xFFFFu));
  n.flags = p.nzFlags >> 16;
  n.tri = p.tri;
  n.b1 = p.b1; n.b2 = p.b2;
  n.pdfFwd = p.pdfFwd; n.pdfBwd = p.pdfBwd;
  n.iorPrev = p.iorPrev; n.iorCur = p.iorCur;
  return n;
}

// ---------------------------------------------------------------- BDPTRayState (:74-186)
struct BdptState {
  float ior[MAX_VOLUMES];
  Spectrum absorption[MAX_VOLUMES];
  uint32_t mat[MAX_VOLUMES];
  uint32_t n;
  uint32_t flags;  // RS_SECONDARY_TERMINATED | RS_REGULARIZE
  YPT_HD void init() {
    n = 0;
    flags = 0;
    for (int i = 0; i < MAX_VOLUMES; ++i) { ior[i] = 1.f; absorption[i].set(0.f); mat[i] = VOLUME_NONE; }
  }
  YPT_HD Spectrum get_absorption() const { return n != 0 ? sel_get(absorption, n - 1) : Spectrum(0.f); }
  YPT_HD float current_ior() const { return n != 0 ? sel_get(ior, n - 1) : 1.0f; }
  YPT_HD float previous_ior() const { return n > 1 ? sel_get(ior, n - 2) : 1.0f; }
  YPT_HD void entered(float i, const Spectrum& a, uint32_t m) {
    n = n + 1 < (uint32_t)MAX_VOLUMES ? n + 1 : (uint32_t)MAX_VOLUMES;
    sel_set(ior, n - 1, i);
    sel_set(absorption, n - 1, a);
    sel_set(mat, n - 1, m);
  }
  // removes the innermost entry of material m (the innermost entry when m is not on the stack)
  YPT_HD void exited(uint32_t m) {
    if (n == 0) return;
    uint32_t index = n - 1;
#pragma unroll
    for (int i = 0; i < MAX_VOLUMES; ++i) index = ((uint32_t)i < n && mat[i] == m) ? (uint32_t)i : index;
#pragma unroll
    for (int i = 0; i + 1 < MAX_VOLUMES; ++i) {
      bool shift = (uint32_t)i >= index && (uint32_t)i + 1 < n;
      ior[i] = pick(shift, ior[i + 1], ior[i]);
      absorption[i] = pick(shift, absorption[i + 1], absorption[i]);
      mat[i] = pick(shift, mat[i + 1], mat[i]);
    }
    --n;
  }
};

// ---------------------------------------------------------------- surfaces at path vertices
struct TriGeometry {
  vec3 p1, p2, p3;
  vec3 geometryNormal, normal, tangent;
};

// positions, geometry normal, interpolated shading normal and tangent (fillSurfaceDefRGB's fetches)
YPT_HD void fetch_triangle(const SceneView& S, uint32_t tri, float b1, float b2, TriGeometry& g) {
  float b0 = 1.f - b1 - b2;
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  g.p1 = load3(S.pos, i0); g.p2 = load3(S.pos, i1); g.p3 = load3(S.pos, i2);
  g.geometryNormal = normalize(cross(g.p2 - g.p1, g.p3 - g.p1));
  g.normal = normalize(b0 * load3(S.nrm, i0) + b1 * load3(S.nrm, i1) + b2 * load3(S.nrm, i2));
  g.tangent = vec3(1.f, 0.f, 0.f);
  if (S.tan) {
    const float* t0 = S.tan + 4 * i0;
    const float* t1 = S.tan + 4 * i1;
    const float* t2 = S.tan + 4 * i2;
    vec3 a(t0[0] * t0[3], t0[1] * t0[3], t0[2] * t0[3]);
    vec3 b(t1[0] * t1[3], t1[1] * t1[3], t1[2] * t1[3]);
    vec3 c(t2[0] * t2[3], t2[1] * t2[3], t2[2] * t2[3]);
    g.tangent = normalize(b0 * a + b1 * b + b2 * c);
  }
}

struct BdptVertex {
  SurfaceContext ctx;
  vec3 pos, wo, normalWS;
  uint32_t tri, matId;
  float b1, b2, t;
  bool frontFace;
};

// makePathHit + calculateCommonSurfaceParams (BDPTCommon.hlsl:490-505, 637-665): the surface hit by
// a ray, two-sided orientation, regularisation, medium entry of rays that start inside, and the
// sampling context (reciprocal energy compensation).
YPT_HD void bdpt_surface(const SceneView& S, BdptState& rs, const Hit& h, vec3 rayOrigin, vec3 rayDir, const Lambdas& L,
                         BdptVertex& v, const MaterialGPU** matOut) {
  TriGeometry g;
  fetch_triangle(S, h.tri, h.b1, h.b2, g);
  uint32_t matId = S.inst[S.triInst[h.tri]].material;
  const MaterialGPU& m = S.mats[matId];
  v.tri = h.tri;
  v.matId = matId;
  v.b1 = h.b1;
  v.b2 = h.b2;
  v.t = h.t;
  v.pos = rayOrigin + rayDir * h.t;
  v.wo = normalize(-rayDir);
  v.frontFace = dot(rayDir, cross(g.p2 - g.p1, g.p3 - g.p1)) < 0.f;
  bool reg = (rs.flags & RS_REGULARIZE) != 0;
  SurfaceTexels tx;
  sample_surface_texels(S, m, h.tri, h.b1, h.b2, tx);
  SurfaceDef sd = make_surface(m, g.geometryNormal, g.normal, g.tangent, v.wo, true, reg, L, S.spec, &tx);
  // handleRayStartInsideMedium (hitShadersCommon.hlsl:213-222)
  if (dot(-v.wo, sd.geometryNormal) > 0.f && rs.n == 0 && !surface_two_sided(sd.flags))
    rs.entered(sd.dielectricIOR, sd.absorption, matId);
  make_surface_context(v.ctx, sd, rs.current_ior(), rs.previous_ior(), v.wo, v.frontFace, L, true);
  v.normalWS = normalize(sd.geometryNormal);
  if (matOut) *matOut = &m;
}

// makePathHitFromSubsurfaceExit + calculateCommonSurfaceParams (BDPTCommon.hlsl:507-521, 637-655):
// the surface at a subsurface exit point, set up with the outward normal as the incoming direction
YPT_HD void bdpt_surface_exit(const SceneView& S, BdptState& rs, const SurfacePoint& p, const Lambdas& L, BdptVertex& v,
                              const MaterialGPU** matOut) {
  TriGeometry g;
  fetch_triangle(S, p.tri, p.b1, p.b2, g);
  uint32_t matId = S.inst[S.triInst[p.tri]].material;
  const MaterialGPU& m = S.mats[matId];
  v.tri = p.tri;
  v.matId = matId;
  v.b1 = p.b1;
  v.b2 = p.b2;
  v.t = 0.f;
  v.pos = p.pos;
  v.wo = normalize(p.n);
  v.frontFace = p.frontFace;
  bool reg = (rs.flags & RS_REGULARIZE) != 0;
  SurfaceTexels tx;
  sample_surface_texels(S, m, p.tri, p.b1, p.b2, tx);
  SurfaceDef sd = make_surface(m, g.geometryNormal, g.normal, g.tangent, v.wo, true, reg, L, S.spec, &tx);
  if (dot(-v.wo, sd.geometryNormal) > 0.f && rs.n == 0 && !surface_two_sided(sd.flags))
    rs.entered(sd.dielectricIOR, sd.absorption, matId);
  make_surface_context(v.ctx, sd, rs.current_ior(), rs.previous_ior(), v.wo, v.frontFace, L, true);
  v.normalWS = normalize(sd.geometryNormal);
  if (matOut) *matOut = &m;
}

YPT_HD Spectrum transmittance(float t, const Spectrum& a) {
  Spectrum tr;
  for (int i = 0; i < 4; ++i) tr.s[i] = yexp(-a.s[i] * t);
  return tr;
}

struct ScatterEval {
  Spectrum weight;
  float pdfFwd, pdfBwd;
  bool dispersed;
};

// evaluateSurfaceAndPDFs (BDPTCommon.hlsl:386-442): forward weight and pdf, the pdf of the reversed
// pair (the surface data recomputed for wi, IORs swapped on transmitted pairs), then the volume
// and dispersion bookkeeping of the event.
YPT_HD ScatterEval eval_and_pdfs(const SceneView& S, const SurfaceContext& ctx, BdptState& rs, uint32_t matId, vec3 wo,
                                 vec3 wi, bool frontFace, bool fromLight, bool treatAsDelta) {
  ScatterEval r;
  r.dispersed = false;
  uint32_t flags = EVAL_NONE;
  if (fromLight) flags |= EVAL_LIGHT_PATH;
  if (treatAsDelta) flags |= EVAL_TREAT_AS_DELTA;
  SurfaceEval fwd = evaluate_surface(ctx, S.luts, wo, wi, flags);
  r.weight = fwd.weight;
  r.pdfFwd = fwd.pdf;
  {
    const SurfaceDef& sd = ctx.def;
    bool transmittedPair = dot(wo, sd.geometryNormal) * dot(wi, sd.geometryNormal) < 0.f;
    bool frontRev = transmittedPair ? !frontFace : frontFace;
    SurfaceContext rev;
    make_surface_context(rev, sd, rs.current_ior(), rs.previous_ior(), wi, frontRev, ctx.lambdas, true);
    if (transmittedPair) {
      rev.pre.fromIOR = ctx.pre.toIOR;
      rev.pre.toIOR = ctx.pre.fromIOR;
    }
    r.pdfBwd = evaluate_surface(rev, S.luts, wi, wo, EVAL_PDF_ONLY).pdf;
  }
  if (fwd.transmission != TRANSMISSION_NONE) {
    if ((fwd.transmission & TRANSMISSION_DISPERSED) != 0) {
      r.dispersed = (rs.flags & RS_SECONDARY_TERMINATED) == 0;
      rs.flags |= RS_SECONDARY_TERMINATED;
    }
    if ((fwd.transmission & TRANSMISSION_EXITED) != 0) rs.exited(matId);
    else if ((fwd.transmission & TRANSMISSION_ENTERED) != 0) rs.entered(ctx.def.dielectricIOR, ctx.def.absorption, matId);
  }
  return r;
}

// sampleScatteringEvent (scatteringSample.hlsl:30-63): a local direction, a subsurface entry (with
// its sampled exit point), the exit lobe at a subsurface exit, or nothing
static constexpr int EV_NONE = 0, EV_LOCAL = 1, EV_SSS_ENTRY = 2, EV_SSS_EXIT_LOBE = 3;
struct ScatterEvent {
  SubsurfaceExit exit;  // EV_SSS_ENTRY
  vec3 wi;              // EV_LOCAL, EV_SSS_EXIT_LOBE
  int type;
  bool isDiffuse, isDelta;
};

template <class Tracer>
YPT_HD void sample_scattering_event(const SceneView& S, const BdptVertex& v, bool atExit, const float r[4], Tracer& tr,
                                    ScatterEvent& ev) {
  ev.type = EV_NONE;
  ev.wi = vec3(0.f);
  ev.isDiffuse = true;
  ev.isDelta = false;
  if (atExit) {
    ev.wi = sample_subsurface_exit_lobe(v.ctx, v.wo, r[2], r[3]);
    if (!is_zero(ev.wi)) ev.type = EV_SSS_EXIT_LOBE;
    return;
  }
  int layer = select_layer(r[0], v.ctx.prob);
  if (layer == LAYER_SUBSURFACE) {
    SurfacePoint e;
    e.pos = v.pos;
    e.n = orient_to_same_side(v.wo, v.normalWS);  // makeSurfacePoint
    e.inst = S.triInst[v.tri];
    e.tri = v.tri;
    e.b1 = v.b1;
    e.b2 = v.b2;
    e.frontFace = v.frontFace;
    Spectrum radius = spectrum_unbounded_no_illum(S.mats[v.matId].subsurfaceRadius, v.ctx.lambdas);
    if (sample_subsurface_exit(S, v.ctx.def.albedo, radius, e, r[1], r[2], r[3], tr, ev.exit)) ev.type = EV_SSS_ENTRY;
    return;
  }
  ev.wi = sample_layer_direction(v.ctx, S.luts, layer, r[1], r[2], r[3], ev.isDiffuse, ev.isDelta);
  if (!is_zero(ev.wi)) ev.type = EV_LOCAL;
}

// evaluateSubsurfaceEntry (scatteringSample.hlsl:66-76): the forward density is one of area (the
// subsurface edge), the backward one that of the exit lobe leaving along wo
YPT_HD ScatterEval eval_subsurface_entry(const SceneView& S, const BdptVertex& v, const SubsurfaceExit& ex) {
  ScatterEval r;
  r.weight = ex.weight * (subsurface_entry_factor(v.ctx, S.luts) * ex.pdfPosition);
  r.pdfFwd = v.ctx.prob[LAYER_SUBSURFACE] * ex.pdfPosition;
  r.dispersed = false;
  Spectrum unused;
  evaluate_subsurface_exit_lobe(v.ctx, v.wo, v.wo, EVAL_PDF_ONLY, unused, r.pdfBwd);
  return r;
}

// evaluateSubsurfaceExitAndPDFs (BDPTCommon.hlsl:463-472): the exit lobe forward, and backward the
// area density of the entry point seen from the exit, with the reversed surface's subsurface
// probability
YPT_HD ScatterEval eval_subsurface_exit(const SceneView& S, const SurfaceContext& ctx, const BdptState& rs, uint32_t matId,
                                        vec3 wo, vec3 wi, bool frontFace, bool fromLight, vec3 exitPos, vec3 exitN,
                                        vec3 entryPos, vec3 entryN) {
  ScatterEval r;
  r.dispersed = false;
  evaluate_subsurface_exit_lobe(ctx, wo, wi, fromLight ? EVAL_LIGHT_PATH : EVAL_NONE, r.weight, r.pdfFwd);
  SurfaceContext rev;
  make_surface_context(rev, ctx.def, rs.current_ior(), rs.previous_ior(), wi, frontFace, ctx.lambdas, true);
  Spectrum radius = spectrum_unbounded_no_illum(S.mats[matId].subsurfaceRadius, ctx.lambdas);
  r.pdfBwd = rev.prob[LAYER_SUBSURFACE] * pdf_subsurface_exit(ctx.def.albedo, radius, exitPos, exitN, entryPos, entryN);
  return r;
}

// ---------------------------------------------------------------- MIS bookkeeping
YPT_HD float pdf_ratio(float nom, float denom) {  // calculatePDFRatio (:548-553)
  nom = nom < 0.f ? 1.f : nom;
  denom = denom < 0.f ? 1.f : denom;
  return fmin_(safe_div(nom, denom), BDPT_MAX_PDF_RATIO_SUM);
}
YPT_HD float accumulate_ri(float riSum, float nom, float denom, bool connectable = true) {  // accumulateRISum (:555-566)
  float ri = pdf_ratio(nom, denom);
  bool acc = connectable && nom > 0.f && denom > 0.f;
  riSum = ri * riSum;
  if (acc) riSum += ri;
  return fmin_(riSum, BDPT_MAX_PDF_RATIO_SUM);
}

// calculateCameraDirectionalPDF (:527-542): directional density of the camera toward a point, with
// one cosine left out (it cancels against the camera's geometry term).
YPT_HD float camera_dir_pdf(const FrameConstants& fc, vec3 camToNodeWS) {
  float a[4] = {0.f, 0.f, 0.f, 1.f}, b[4] = {1.f, 1.f, 0.f, 1.f}, mn[4], mx[4];
  mul4(fc.uvToView, a, mn);
  mul4(fc.uvToView, b, mx);
  float wn = mn[3], wx = mx[3];
  for (int i = 0; i < 4; ++i) { mn[i] = mn[i] / wn; mx[i] = mx[i] / wx; }
  float ex = fabsf(mx[0] - mn[0]), ey = fabsf(mx[1] - mn[1]);
  float planeArea = safe_div(ex * ey, mn[2] * mn[2]);
  vec3 vs = xform_dir(fc.worldToView, camToNodeWS);
  float cosTheta = fabsf(vs.z) * (1.f / sqrtf(fmax_(0.0001f, dot(vs, vs))));
  return safe_div(fc.targetTexDimensions[0] * fc.targetTexDimensions[1], planeArea * cosTheta * cosTheta * cosTheta);
}

// pdfForSamplingLight (lightSampling.hlsl:191-228) for a point on a mesh-light triangle
YPT_HD void pdf_sampling_light(const SceneView& S, uint32_t tri, vec3 lightNormal, vec3 towards, float& pick, float& posPdf,
                               float& dirPdf) {
  const InstanceGPU& in = S.inst[S.triInst[tri]];
  const MaterialGPU& m = S.mats[in.material];
  pick = (1.f - 0.f) * in.lightSelectionPdf;  // (1 - environmentSelectionPDF) with mesh lights present
  bool twoSided = surface_two_sided(m.flags);
  vec3 p1, p2, p3;
  tri_vertices(S, tri, p1, p2, p3);
  float area = length(cross(p2 - p1, p3 - p1)) * 0.5f;
  posPdf = 1.f / ((float)in.triCount * area);
  float cosTheta = dot(lightNormal, towards);
  if (twoSided) dirPdf = pdf_cosine_power_sphere(cosTheta, m.emissionFocus);
  else dirPdf = cosTheta >= 0.f ? pdf_cosine_power_hemisphere(cosTheta, m.emissionFocus) : 0.f;
}

// pdfEnvironmentPosition (lightSampling.hlsl:73-76): planar density of the disk behind the world bounds
YPT_HD float pdf_environment_position(const FrameConstants& fc) {
  vec3 lo(fc.worldBoundsMin[0], fc.worldBoundsMin[1], fc.worldBoundsMin[2]);
  vec3 hi(fc.worldBoundsMax[0], fc.worldBoundsMax[1], fc.worldBoundsMax[2]);
  vec3 c = (hi + lo) * 0.5f;
  vec3 ext = hi - c;
  return safe_div(1.f, PI * dot(ext, ext));
}

// getWorldCenterAndRadiusSqr (raytraceCommonResources.hlsl:99-104)
YPT_HD vec3 world_center_radius_sqr(const FrameConstants& fc, float& radiusSqr) {
  vec3 lo(fc.worldBoundsMin[0], fc.worldBoundsMin[1], fc.worldBoundsMin[2]);
  vec3 hi(fc.worldBoundsMax[0], fc.worldBoundsMax[1], fc.worldBoundsMax[2]);
  vec3 c = (hi + lo) * 0.5f;
  vec3 ext = hi - c;
  radiusSqr = dot(ext, ext);
  return c;
}

// environmentSelectionPDF (lightSampling.hlsl:25-28): light subpaths start from the environment only
// when the scene has no lights
YPT_HD float environment_selection_pdf(const FrameConstants& fc) { return fc.lightCount == 0 ? 1.f : 0.f; }

// pdfEnvironmentDirection (lightSampling.hlsl:64-70): directions uniform in spherical coordinates about +z
YPT_HD float pdf_environment_direction(vec3 dir) {
  float cosTheta = clampf(dir.z, -1.f, 1.f);
  float sinTheta = sqrtf(fmax_(0.f, 1.f - cosTheta * cosTheta));
  return safe_div(1.f, 2.f * PI * PI * sinTheta);
}

// shouldPathContribute (BDPTCommon.hlsl:1217-1223): the renderer's debug filter of strategies by their
// camera and light vertex counts (_DEBUG.BDPT.EffectiveCameraPathNodes / EffectiveLightPathNodes)
YPT_HD bool should_contribute(const FrameConstants& fc, uint32_t cameraNodes, uint32_t lightNodes) {
  return cameraNodes >= fc.nodeRanges[2] && cameraNodes <= fc.nodeRanges[3] && lightNodes >= fc.nodeRanges[0] &&
         lightNodes <= fc.nodeRanges[1];
}

// getSkyBoxColor times the scale (as the miss shader sees it); the white furnace's unit radiance
YPT_HD vec3 sky_color(const SceneView& S, const FrameConstants& fc, vec3 dir) {
  if (fc.flags & FRAME_FLAG_WHITE_FURNACE) return vec3(1.f);
  return env_color(S, dir) * fc.envIntensityScale;
}

// encodeOctahedral / decodeOctahedral (commonMath.hlsl:96-117)
YPT_HD void encode_octahedral(vec3 n, float& px, float& py) {
  n = n / (fabsf(n.x) + fabsf(n.y) + fabsf(n.z));
  px = n.x; py = n.y;
  if (n.z < 0.f) {
    float sx = px >= 0.f ? 1.f : -1.f, sy = py >= 0.f ? 1.f : -1.f;
    float ax = (1.f - fabsf(py)) * sx, ay = (1.f - fabsf(px)) * sy;
    px = ax; py = ay;
  }
}
YPT_HD vec3 decode_octahedral(float px, float py) {
  vec3 n(px, py, 1.f - fabsf(px) - fabsf(py));
  if (n.z < 0.f) {
    float sx = n.x >= 0.f ? 1.f : -1.f, sy = n.y >= 0.f ? 1.f : -1.f;
    float ax = (1.f - fabsf(n.y)) * sx, ay = (1.f - fabsf(n.x)) * sy;
    n.x = ax; n.y = ay;
  }
  return normalize(n);
}

// sampleSphere, sampleUniformCone, pdfUniformCone (miscBrdf.hlsl:70-94)
YPT_HD vec3 sample_sphere(float u0, float u1) {
  float y = 1.f - 2.f * u0;
  float r = sqrtf(1.f - sqr(y));
  float s, c;
  ysincos(2.f * PI * u1, &s, &c);
  return vec3(r * c, y, r * s);
}
YPT_HD vec3 sample_uniform_cone(float u0, float u1, float cosMax) {
  float y = 1.f - u0 * (1.f - cosMax);
  float r = sqrtf(fmax_(0.f, 1.f - sqr(y)));
  float s, c;
  ysincos(2.f * PI * u1, &s, &c);
  return vec3(r * c, y, r * s);
}
YPT_HD float pdf_uniform_cone(float cosMax) { return safe_div(1.f, 2.f * PI * (1.f - cosMax)); }
YPT_HD float pdf_sphere() { return 1.f / (4.f * PI); }

// pdfPunctualLightEmissionDirection (lightSampling.hlsl:294-307): 1 for directional lights (delta)
YPT_HD float pdf_punctual_emission_direction(const LightGPU& l, vec3 dirFromLight) {
  if (l.type == LIGHT_TYPE_SPOT) {
    vec3 axis(l.directionCosOuter[0], l.directionCosOuter[1], l.directionCosOuter[2]);
    return dot(dirFromLight, axis) >= l.directionCosOuter[3] ? pdf_uniform_cone(l.directionCosOuter[3]) : 0.f;
  }
  if (l.type == LIGHT_TYPE_POINT) return pdf_sphere();
  return 1.f;
}

// direction in which an infinite light origin (directional light or environment) emits
// (getInfiniteLightNodeDirection, BDPTCommon.hlsl:600-607); environment nodes keep it octahedral
// encoded in their barycentrics
YPT_HD bool infinite_node(uint32_t flags) { return (flags & (LN_DIRECTIONAL | LN_ENV)) != 0; }
YPT_HD vec3 infinite_node_direction(const SceneView& S, const LightNode& n) {
  if (n.flags & LN_DIRECTIONAL) {
    const LightGPU& l = S.lights[n.tri];
    return vec3(l.directionCosOuter[0], l.directionCosOuter[1], l.directionCosOuter[2]);
  }
  return decode_octahedral(n.b1, n.b2);
}

// ---------------------------------------------------------------- light subpath
YPT_HD uint32_t camera_dim_offset(const FrameConstants& fc) { return fc.lightPathVertices * 5u + 12u; }      // :226-227
YPT_HD uint32_t pure_lens_dim_offset(const FrameConstants& fc) { return fc.lightPathVertices * 5u + 9u; }    // :229

// traceLightRays (BDPTTraceLightPath.comp:6-362) for light subpath `pathIndex`: writes the vertices,
// light first, to out[0 .. count) and returns count (0 when the light sample failed).
template <class Tracer>
YPT_HD uint32_t trace_light_path(const SceneView& S, const FrameConstants& fc, const uint32_t* point, uint32_t pathIndex,
                                 Tracer& tr, LightNodePacked* out, float& lambdaOut) {
  RandomSampler rand;
  rand.init(point, 0u, pathIndex);
  float lambda = hero_wavelength(rand.r1());
  lambdaOut = lambda;
  Lambdas L = spectral_lambdas(lambda);
  float r0[4], r1[4];
  rand.r4(r0);
  rand.r4(r1);
  // sampleLightOrEnv (BDPTCommon.hlsl:680-720): the environment when the scene has no lights, else a
  // light by the alias table, sampled by sampleLight (mesh) or samplePunctualLightEmission
  float envPdf = environment_selection_pdf(fc);
  Spectrum radiance;
  vec3 lightPos, lightDir, lightNormal;
  float pdfPos, pdfDir, pick, profile;
  uint32_t originTri = NO_HIT, originFlags = 0;
  float originB1 = -1.f, originB2 = -1.f;
  if (r1[3] < envPdf) {
    // sampleEnvironmentLighting (lightSampling.hlsl:78-120)
    float st, ct, sp, cp;
    ysincos(r0[0] * PI, &st, &ct);
    ysincos(r0[1] * 2.f * PI, &sp, &cp);
    lightDir = vec3(st * cp, st * sp, ct);
    pdfDir = safe_div(1.f, 2.f * PI * PI * st);
    float rsq;
    vec3 c = world_center_radius_sqr(fc, rsq);
    vec3 v1, v2;
    construct_vector_base(lightDir, v1, v2);
    float cx, cy;
    sample_concentric_disk(r0[2], r0[3], cx, cy);
    float radius = sqrtf(rsq);
    lightPos = c - lightDir * radius + radius * (cx * v1 + cy * v2);
    pdfPos = safe_div(1.f, PI * rsq);
    radiance = rgb_unbounded_spectrum(S, sky_color(S, fc, -lightDir), L);
    lightNormal = lightDir;  // the disk faces the emitted direction
    profile = 1.f;
    pick = envPdf;
    originFlags = LN_ENV;
    encode_octahedral(lightDir, originB1, originB2);
  } else {
    uint32_t li = sample_light_index(S.lights, fc.lightCount, (r1[3] - envPdf) / (1.f - envPdf), pick);
    pick = (1.f - envPdf) * pick;
    const LightGPU& light = S.lights[li];
    if (light.type != LIGHT_TYPE_MESH) {
      // samplePunctualLightEmission (lightSampling.hlsl:311-357) with random values (r1.x, r1.y, r0.x, r0.y)
      radiance = rgb_unbounded_spectrum(S, vec3(light.intensity[0], light.intensity[1], light.intensity[2]), L);
      originTri = li;
      originFlags = LN_PUNCTUAL;
      vec3 axis(light.directionCosOuter[0], light.directionCosOuter[1], light.directionCosOuter[2]);
      if (light.type == LIGHT_TYPE_DIRECTIONAL) {
        originFlags |= LN_DIRECTIONAL;
        float rsq;
        vec3 c = world_center_radius_sqr(fc, rsq);
        float radius = sqrtf(rsq);
        vec3 v1, v2;
        construct_vector_base(axis, v1, v2);
        float cx, cy;
        sample_concentric_disk(r1[0], r1[1], cx, cy);
        lightPos = c - axis * radius + radius * (cx * v1 + cy * v2);
        lightDir = axis;
        pdfPos = safe_div(1.f, PI * rsq);  // pdfDirectionalLightPosition
        pdfDir = 1.f;
        profile = 1.f;
      } else {
        lightPos = vec3(light.positionCosInner[0], light.positionCosInner[1], light.positionCosInner[2]);
        pdfPos = 1.f;
        if (light.type == LIGHT_TYPE_SPOT) {
          vec3 v1, v2;
          construct_vector_base(axis, v1, v2);
          vec3 local = sample_uniform_cone(r1[0], r1[1], light.directionCosOuter[3]);
          lightDir = normalize(local.x * v1 + local.y * axis + local.z * v2);
          pdfDir = pdf_uniform_cone(light.directionCosOuter[3]);
          profile = punctual_falloff(light, lightDir);
        } else {
          lightDir = sample_sphere(r1[0], r1[1]);
          pdfDir = pdf_sphere();
          profile = 1.f;
        }
      }
      lightNormal = lightDir;  // the cosine at the light is 1
    } else {
      // sampleLight (lightSampling.hlsl:122-189)
      uint32_t primCount = light.triCount;
      uint32_t prim = (uint32_t)fmin_((float)primCount * r0[2], (float)(primCount - 1u));
      float bx = r0[0], by = r0[1];
      if (bx + by > 1.f) { bx = 1.f - bx; by = 1.f - by; }
      uint32_t tri = light.firstTri + prim;
      TriGeometry g;
      fetch_triangle(S, tri, bx, by, g);
      vec3 gn = cross(g.p2 - g.p1, g.p3 - g.p1);
      float gl = length(gn);
      gn = gn / gl;
      float area = gl * 0.5f;
      float b0 = 1.f - bx - by;
      lightPos = b0 * g.p1 + bx * g.p2 + by * g.p3;
      vec3 tangent = normalize(g.tangent);
      const MaterialGPU& lm = S.mats[light.material];
      bool twoSided = surface_two_sided(lm.flags);
      float focus = lm.emissionFocus;
      vec3 local = twoSided ? sample_cosine_power_sphere(r1[0], r1[1], focus) : sample_cosine_power_hemisphere(r1[0], r1[1], focus);
      mat3 tanToWS = basis_transform(gn, tangent);
      radiance = surface_emission(S, lm, tri, bx, by, L);
      if (alpha_cut(S, tri, bx, by)) radiance.set(0.f);
      pdfPos = 1.f / ((float)primCount * area);
      pdfDir = twoSided ? pdf_cosine_power_sphere(local.y, focus) : pdf_cosine_power_hemisphere(local.y, focus);
      profile = emission_profile(local.y, focus);
      lightDir = mul(local, tanToWS);
      lightNormal = gn;
      originTri = tri;
    }
  }
  if (radiance.all_equal(0.f) || pdfPos == 0.f || pdfDir == 0.f || pick == 0.f) return 0;
  float maxRadiance = radiance.max_value();
  const bool originPunctual = (originFlags & LN_PUNCTUAL) != 0;
  const bool originInfinite = infinite_node(originFlags);
  const bool originEnv = (originFlags & LN_ENV) != 0;

  LightNode nodes[2];
  nodes[0].radiance = radiance;
  nodes[0].pos = lightPos;
  nodes[0].nrm = lightNormal;
  nodes[0].tri = originTri;
  nodes[0].b1 = originB1;
  nodes[0].b2 = originB2;
  // for the environment the origin is the direction: its density is the directional one
  nodes[0].pdfFwd = originEnv ? pdfDir * pick : pdfPos * pick;
  nodes[0].pdfBwd = 0.f;
  nodes[0].iorPrev = 1.f;
  nodes[0].iorCur = 1.f;
  nodes[0].flags = originFlags;
  nodes[1] = nodes[0];
  nodes[1].pdfBwd = 0.f;
  float pdfForward = pdfDir;
  vec3 rayDir = lightDir;
  vec3 rayOrigin = spawn_ray_origin(lightPos, lightNormal, lightDir);
  radiance = radiance * profile * fabsf(dot(lightNormal, lightDir)) / (pick * pdfPos * pdfDir);

  uint32_t stored = 0, nodeCount = 1;
  bool previousWasDelta = false;
  BdptState rs;
  rs.init();
  // set when the previous vertex took the subsurface event: the next vertex is its exit point
  bool pendingExit = false;
  SubsurfaceExit exit;
  for (uint32_t i = 1; i < fc.lightPathVertices; ++i) {
    float rs0[4];
    rand.r4(rs0);
    Hit h;
    bool isExit = pendingExit;
    pendingExit = false;
    if (!isExit && !tr.closest(rayOrigin, rayDir, RAY_MIN_T, __builtin_inff(), h)) break;
    uint32_t ci = i & 1u, pi = (i - 1u) & 1u;
    if (i - 1u >= fc.regularizeAfterVertices) rs.flags |= RS_REGULARIZE;
    uint32_t flags = 0;
    BdptVertex v;
    if (isExit) bdpt_surface_exit(S, rs, exit.p, L, v, nullptr);
    else bdpt_surface(S, rs, h, rayOrigin, rayDir, L, v, nullptr);
    if (v.frontFace) flags |= LN_FRONT_FACE;
    if (rs.flags & RS_SECONDARY_TERMINATED) flags |= LN_SECONDARY_TERMINATED;
    if (rs.flags & RS_REGULARIZE) flags |= LN_REGULARIZED;
    if (isExit) flags |= LN_SUBSURFACE_EXIT;
    // a subsurface exit only has the exit lobe
    else if (is_delta_distribution(v.ctx.pre.ax, v.ctx.pre.ay, v.ctx.def)) flags |= LN_DELTA;
    vec3 nrm = isExit ? orient_to_same_side(v.wo, v.normalWS) : v.normalWS;  // exit nodes keep the outward normal
    float curFwd;
    if (isExit) {
      curFwd = pdfForward;  // the subsurface edge density is one of area already
    } else {
      if (i == 1u && originInfinite) curFwd = pdfPos * fabsf(dot(rayDir, v.normalWS));  // planar density of the disk
      else curFwd = previousWasDelta ? -1.f : pdfForward * solid_angle_to_area(v.pos - rayOrigin, v.normalWS);
      Spectrum ab = rs.get_absorption();
      if (!ab.all_equal(0.f)) radiance = radiance * transmittance(h.t, ab);
    }
    nodes[ci].radiance = radiance;
    nodes[ci].pos = v.pos;
    nodes[ci].nrm = nrm;
    nodes[ci].tri = v.tri;
    nodes[ci].b1 = v.b1;
    nodes[ci].b2 = v.b2;
    nodes[ci].pdfFwd = curFwd;
    nodes[ci].iorPrev = rs.previous_ior();
    nodes[ci].iorCur = rs.current_ior();
    nodes[ci].flags = flags;

    ScatterEvent se;
    sample_scattering_event(S, v, isExit, rs0, tr, se);
    if (se.type == EV_NONE) break;
    if (se.isDiffuse) rs.flags |= RS_REGULARIZE;
    ScatterEval ev;
    bool bwdIsArea = false;
    if (se.type == EV_SSS_ENTRY) {
      ev = eval_subsurface_entry(S, v, se.exit);
    } else if (se.type == EV_SSS_EXIT_LOBE) {
      ev = eval_subsurface_exit(S, v.ctx, rs, v.matId, v.wo, se.wi, v.frontFace, true, v.pos, nrm, nodes[pi].pos, nodes[pi].nrm);
      bwdIsArea = true;
    } else {
      ev = eval_and_pdfs(S, v.ctx, rs, v.matId, v.wo, se.wi, v.frontFace, true, se.isDelta);
    }
    pdfForward = ev.pdfFwd;
    Spectrum sampleWeight = ev.weight * safe_div(1.f, pdfForward);
    nodes[pi].pdfBwd = se.isDelta ? -1.f : (bwdIsArea ? ev.pdfBwd : ev.pdfBwd * solid_angle_to_area(rayOrigin - v.pos, nodes[pi].nrm));
    if (i == 1u && originPunctual) nodes[pi].pdfBwd = 0.f;                    // a camera path never hits a delta light
    else if (i == 1u && originEnv && !se.isDelta) nodes[pi].pdfBwd = ev.pdfBwd;  // found by direction: stays per solid angle
    out[stored++] = pack_node(nodes[pi]);
    ++nodeCount;
    radiance = radiance * sampleWeight;
    if (ev.dispersed) radiance.terminate_secondary();
    if (radiance.russian_roulette(rand.r1(), maxRadiance)) break;
    if (radiance.all_equal(0.f)) break;
    if (se.type == EV_SSS_ENTRY) {
      pendingExit = true;
      exit = se.exit;
    } else {
      rayDir = normalize(se.wi);
      rayOrigin = spawn_ray_origin(v.pos, nrm, rayDir);
    }
    previousWasDelta = se.isDelta;
  }
  out[stored++] = pack_node(nodes[(nodeCount - 1u) & 1u]);
  return stored;
}

// ---------------------------------------------------------------- connections
// The surface of a light vertex as a connection sees it: material, shading frame and layer
// probabilities, with wo towards the vertex before it. None of it depends on the connection's
// other end, so it can be made once per light vertex and shared by every connection to it.
struct LightNodeSurface {
  SurfaceContext ctx;
  vec3 wo;
};

YPT_HD BdptState light_node_state(const LightNode& n, bool secondaryTerminated) {
  BdptState rs;
  rs.init();
  rs.ior[0] = n.iorPrev;
  rs.ior[1] = n.iorCur;
  rs.n = 2;
  bool already = secondaryTerminated || (n.flags & LN_SECONDARY_TERMINATED) != 0;
  rs.flags = already ? RS_SECONDARY_TERMINATED : 0u;
  return rs;
}

// light vertex `n` (index > 0) reached from prevPos
YPT_HD void light_node_surface(const SceneView& S, const LightNode& n, vec3 prevPos, const Lambdas& L, LightNodeSurface& out) {
  const MaterialGPU& m = S.mats[S.inst[S.triInst[n.tri]].material];
  bool frontFace = (n.flags & LN_FRONT_FACE) != 0;
  // a subsurface exit node stores the outward normal, which stands in for wo and sets the surface up
  // for the outside (the two-sided orientation applies there); other nodes face the vertex before
  // them, with no two-sided re-orientation, as in the reference
  bool exitNode = (n.flags & LN_SUBSURFACE_EXIT) != 0;
  out.wo = exitNode ? normalize(n.nrm) : normalize(prevPos - n.pos);
  TriGeometry g;
  fetch_triangle(S, n.tri, n.b1, n.b2, g);
  SurfaceTexels tx;
  sample_surface_texels(S, m, n.tri, n.b1, n.b2, tx);
  SurfaceDef sd = make_surface(m, g.geometryNormal, g.normal, g.tangent, out.wo, exitNode, (n.flags & LN_REGULARIZED) != 0, L,
                               S.spec, &tx);
  BdptState rs = light_node_state(n, false);
  make_surface_context(out.ctx, sd, rs.current_ior(), rs.previous_ior(), out.wo, frontFace, L, true);
}

// Sources of the light-vertex surfaces: made for each connection, as the reference does, or read
// from a table made once per vertex (k_bdpt_node_surfaces). Both give the same bits.
struct NodeSurfacesOnTheFly {
  LightNodeSurface tmp;
  YPT_HD const LightNodeSurface* get(const SceneView& S, const LightNode& n, uint32_t, vec3 prevPos, const Lambdas& L) {
    light_node_surface(S, n, prevPos, L, tmp);
    return &tmp;
  }
};
struct NodeSurfaceTable {
  const LightNodeSurface* p;  // [vertex index] of one light subpath
  YPT_HD const LightNodeSurface* get(const SceneView&, const LightNode&, uint32_t index, vec3, const Lambdas&) const {
    return p + index;
  }
};

// calculateConnectingLightNodeWeightsAndPDFs (:701-875) for a mesh-light origin (index 0) or a
// surface vertex with its surface `ls`; false when the connection carries nothing.
YPT_HD bool connect_light_node(const SceneView& S, const FrameConstants& fc, const LightNode& n, uint32_t index,
                               const LightNodeSurface* ls, vec3 nextPos, bool secondaryTerminated, const Lambdas& L,
                               ScatterEval& ev, float& pdfFwdMIS, vec3 prevPos = vec3(0.f), vec3 prevNrm = vec3(0.f)) {
  ev.weight.set(0.f);
  ev.pdfFwd = 0.f;
  ev.pdfBwd = 0.f;
  ev.dispersed = false;
  bool frontFace = (n.flags & LN_FRONT_FACE) != 0;
  if (index == 0 && (n.flags & LN_PUNCTUAL)) {
    // punctual origin: the node radiance is an intensity (point, spot) or irradiance (directional)
    const LightGPU& light = S.lights[n.tri];
    float pick = (1.f - environment_selection_pdf(fc)) * light.selectionPdf;
    if (light.type == LIGHT_TYPE_DIRECTIONAL) {
      ev.pdfFwd = pick;
      ev.weight.set(1.f / pick);
      float rsq;
      world_center_radius_sqr(fc, rsq);
      pdfFwdMIS = safe_div(1.f, PI * rsq);  // planar density reaching the next vertex
    } else {
      vec3 fromLight = nextPos - n.pos;
      float distSqr = fmax_(dot(fromLight, fromLight), 1e-10f);
      vec3 fromLightDir = fromLight * (1.f / sqrtf(distSqr));
      float falloff = punctual_falloff(light, fromLightDir);
      if (falloff <= 0.f) { pdfFwdMIS = 0.f; return false; }
      ev.pdfFwd = pick * distSqr;
      ev.weight.set(falloff / ev.pdfFwd);
      pdfFwdMIS = pdf_punctual_emission_direction(light, fromLightDir);
    }
    ev.pdfBwd = 0.f;
    return true;
  }
  if (index == 0 && (n.flags & LN_ENV)) {
    // environment origin: connecting along its direction is like sampling that direction
    float pdfDir = environment_selection_pdf(fc) * pdf_environment_direction(infinite_node_direction(S, n));
    if (pdfDir <= 0.f) { pdfFwdMIS = 0.f; return false; }
    ev.pdfFwd = pdfDir;
    ev.weight.set(1.f / pdfDir);
    ev.pdfBwd = 0.f;
    pdfFwdMIS = pdf_environment_position(fc);
    return true;
  }
  if (index == 0) {
    const MaterialGPU& m = S.mats[S.inst[S.triInst[n.tri]].material];
    bool twoSided = surface_two_sided(m.flags);
    vec3 fromLight = nextPos - n.pos;
    if (dot(fromLight, n.nrm) <= 0.f && !twoSided) { pdfFwdMIS = 0.f; return false; }
    float pdf = n.pdfFwd;
    ev.pdfFwd = pdf * area_to_solid_angle(-fromLight, n.nrm);
    float cosEmit = dot(n.nrm, normalize(fromLight));
    ev.weight.set(safe_div(emission_profile(cosEmit, m.emissionFocus), ev.pdfFwd));
    pdfFwdMIS = twoSided ? pdf_cosine_power_sphere(cosEmit, m.emissionFocus)
                         : (cosEmit >= 0.f ? pdf_cosine_power_hemisphere(cosEmit, m.emissionFocus) : 0.f);
    return true;
  }
  vec3 wi = normalize(nextPos - n.pos);
  if (n.flags & LN_SUBSURFACE_EXIT) {  // the exit lobe, and the entry point (prevPos) as the backward density
    BdptState rs = light_node_state(n, false);
    rs.flags = 0u;
    ev = eval_subsurface_exit(S, ls->ctx, rs, S.inst[S.triInst[n.tri]].material, ls->wo, wi, frontFace, true, n.pos, n.nrm,
                              prevPos, prevNrm);
    pdfFwdMIS = ev.pdfFwd;
    return !(ev.pdfBwd == 0.f || ev.weight.all_equal(0.f));
  }
  BdptState rs = light_node_state(n, secondaryTerminated);
  ev = eval_and_pdfs(S, ls->ctx, rs, VOLUME_NONE, ls->wo, wi, frontFace, true, false);
  if (ev.dispersed) ev.weight.terminate_secondary();
  pdfFwdMIS = ev.pdfFwd;
  if (ev.pdfBwd == 0.f || ev.weight.all_equal(0.f)) return false;
  return true;
}

struct CameraVertexCtx {  // CameraVertexContext (:877-896)
  Spectrum throughput;
  vec3 pos, normalWS, wo, wi;
  vec3 prevPos, prevNrm;
  float pdfRatioSum;
  float prevFwdMIS, curFwdMIS;
  uint32_t matId;
  bool frontFace;
  bool isExit, prevIsExit;  // reached through the subsurface from the previous vertex (this one, the one before)
};

struct LightVertexCtx {  // LightVertexContext (:898-906)
  vec3 prevPos, prevNrm;
  float prevFwdMIS;
  float riSum;
  uint32_t prevFlags;
  uint32_t originFlags;  // flags of the subpath's light origin (vertex 0)
};

struct ConnectionPdfs {
  float fwdCam, bwdCam, fwdLight, bwdLight;
};

// calculateMISWeight (:918-1007)
YPT_HD float mis_connection(const SceneView& S, const FrameConstants& fc, uint32_t camIndex, const CameraVertexCtx& cv,
                            uint32_t lightIndex, const LightVertexCtx& lv, const LightNode& n, const ConnectionPdfs& p) {
  float riCam = cv.pdfRatioSum;
  if (camIndex > 0) {  // from a subsurface exit the backward density is one of area already
    float pb = cv.isExit ? p.bwdCam : p.bwdCam * solid_angle_to_area(cv.prevPos - cv.pos, cv.prevNrm);
    riCam = accumulate_ri(riCam, pb, cv.prevFwdMIS, !cv.prevIsExit);
  }
  float toCamArea;
  if (lightIndex == 0 && infinite_node(n.flags))  // from a disk: planar density times the cosine
    toCamArea = p.fwdLight * fabsf(dot(cv.normalWS, infinite_node_direction(S, n)));
  else
    toCamArea = p.fwdLight * solid_angle_to_area(cv.pos - n.pos, cv.normalWS);
  riCam = accumulate_ri(riCam, toCamArea, cv.curFwdMIS, !cv.isExit);
  float riLight = lv.riSum;
  if (lightIndex == 0) {
    float pick, posPdf, dirPdf, toLight;
    if (n.flags & LN_PUNCTUAL) {  // a camera path never hits a delta light
      pick = 1.f;
      posPdf = 1.f;
      toLight = 0.f;
    } else if (n.flags & LN_ENV) {  // found by the same direction: both densities per solid angle
      pick = environment_selection_pdf(fc);
      posPdf = pdf_environment_direction(infinite_node_direction(S, n));
      toLight = p.fwdCam;
    } else {
      vec3 camToLight = n.pos - cv.pos;
      float len = length(camToLight);
      vec3 lightToCam = (-camToLight) / len;
      pdf_sampling_light(S, n.tri, n.nrm, lightToCam, pick, posPdf, dirPdf);
      toLight = p.fwdCam * solid_angle_to_area(camToLight, n.nrm);
    }
    riLight = pdf_ratio(toLight, pick * posPdf);
  } else {
    bool prevPunctual = lightIndex == 1 && (lv.originFlags & LN_PUNCTUAL) != 0;
    bool prevEnv = lightIndex == 1 && (lv.originFlags & LN_ENV) != 0;
    bool nodeExit = (n.flags & LN_SUBSURFACE_EXIT) != 0;
    bool prevExit = (lv.prevFlags & LN_SUBSURFACE_EXIT) != 0;
    float pbl = p.bwdLight;
    if (prevPunctual) pbl = 0.f;
    else if (!prevEnv && !nodeExit) pbl = pbl * solid_angle_to_area(lv.prevPos - n.pos, lv.prevNrm);
    riLight = accumulate_ri(riLight, pbl, lv.prevFwdMIS, !prevExit);
    riLight = accumulate_ri(riLight, p.fwdCam * solid_angle_to_area(n.pos - cv.pos, n.nrm), n.pdfFwd, !nodeExit);
  }
  return 1.f / (1.f + riCam + riLight);
}

// calculateMISWeightOnCameraPathHittingLight (:1009-1036)
YPT_HD float mis_camera_hits_light(const SceneView& S, uint32_t camIndex, uint32_t tri, const CameraVertexCtx& cv) {
  if (camIndex == 0) return 1.f;
  vec3 towardsPrev = cv.prevPos - cv.pos;
  vec3 towardsPrevDir = normalize(towardsPrev);
  float pick, posPdf, dirPdf;
  pdf_sampling_light(S, tri, cv.normalWS, towardsPrevDir, pick, posPdf, dirPdf);
  float riCam = cv.pdfRatioSum;
  riCam = accumulate_ri(riCam, dirPdf * solid_angle_to_area(towardsPrev, cv.prevNrm), cv.prevFwdMIS, !cv.prevIsExit);
  riCam = accumulate_ri(riCam, pick * posPdf, cv.curFwdMIS);
  return 1.f / (1.f + riCam);
}

// calculateMISWeightOnCameraPathMissing (:1040-1058)
YPT_HD float mis_camera_misses(const FrameConstants& fc, uint32_t camIndex, const CameraVertexCtx& cv, vec3 rayDir) {
  if (camIndex == 0) return 1.f;
  float riCam = cv.pdfRatioSum;
  float pdfEnvToLast = pdf_environment_position(fc) * fabsf(dot(cv.prevNrm, rayDir));
  riCam = accumulate_ri(riCam, pdfEnvToLast, cv.prevFwdMIS, !cv.prevIsExit);
  float pdfEnvDirection = environment_selection_pdf(fc) * pdf_environment_direction(-rayDir);
  riCam = accumulate_ri(riCam, pdfEnvDirection, cv.curFwdMIS);
  return 1.f / (1.f + riCam);
}

// A connection whose contribution is known but whose visibility is not yet tested: the ray from
// the camera vertex must hit the light vertex's triangle first.
struct PendingConnection {
  Spectrum contrib;
  vec3 origin, dir;
  float tmax;
  uint32_t tri;
};

// calculateLightPathRadiance (:1060-1185) up to its visibility ray; rs is the camera subpath's
// state, taken by value. Returns false when the connection carries nothing. Visibility only gates
// the result, so testing it after the MIS weight changes no bit of the contribution.
YPT_HD bool connection_prepare(const SceneView& S, const FrameConstants& fc, uint32_t camIndex, const CameraVertexCtx& cv,
                               const SurfaceContext& camCtx,
                               const LightNode& n, uint32_t lightIndex, const LightNodeSurface* ls, const LightVertexCtx& lv,
                               BdptState rs, const Lambdas& L, PendingConnection& out) {
  bool camTerminated = (rs.flags & RS_SECONDARY_TERMINATED) != 0;
  bool lightTerminated = (n.flags & LN_SECONDARY_TERMINATED) != 0;
  if (lightTerminated) rs.flags |= RS_SECONDARY_TERMINATED;
  ConnectionPdfs p;
  ScatterEval ec = cv.isExit ? eval_subsurface_exit(S, camCtx, rs, cv.matId, cv.wo, cv.wi, cv.frontFace, false, cv.pos,
                                                    cv.normalWS, cv.prevPos, cv.prevNrm)
                             : eval_and_pdfs(S, camCtx, rs, cv.matId, cv.wo, cv.wi, cv.frontFace, false, false);
  if (ec.dispersed) ec.weight.terminate_secondary();
  if (ec.pdfFwd == 0.f || ec.weight.all_equal(0.f)) return false;
  Spectrum wCam = ec.weight;
  p.fwdCam = ec.pdfFwd;
  p.bwdCam = ec.pdfBwd;
  ScatterEval el;
  if (!connect_light_node(S, fc, n, lightIndex, ls, cv.pos, (rs.flags & RS_SECONDARY_TERMINATED) != 0, L, el, p.fwdLight,
                          lv.prevPos, lv.prevNrm))
    return false;
  Spectrum wLight = el.weight;
  p.bwdLight = el.pdfBwd;
  // getLightNodeConnection (:616-630): infinite origins are reached along their direction; origins
  // without geometry are visible when nothing is hit on the way
  bool atInfinity = lightIndex == 0 && infinite_node(n.flags);
  bool noGeometry = lightIndex == 0 && (n.flags & (LN_PUNCTUAL | LN_ENV)) != 0;
  float len;
  if (atInfinity) {
    out.dir = -infinite_node_direction(S, n);
    len = __builtin_inff();
  } else {
    vec3 toNode = n.pos - cv.pos;
    len = length(toNode);
    out.dir = toNode / len;
  }
  out.origin = spawn_ray_origin(cv.pos, cv.normalWS, out.dir);
  if (noGeometry) {
    out.tmax = atInfinity ? __builtin_inff() : len - RAY_MIN_T;
    out.tri = NO_HIT;
  } else {
    out.tmax = RAY_MIN_T + len + BDPT_VISIBILITY_EPS;
    out.tri = n.tri;
  }
  float mis = mis_connection(S, fc, camIndex, cv, lightIndex, lv, n, p);
  float geom = 1.f;
  if (lightIndex > 0) geom = 1.f / (len * len);
  Spectrum ab = rs.get_absorption();
  if (!atInfinity && !ab.all_equal(0.f)) wCam = wCam * transmittance(len, ab);
  Spectrum r = wLight * wCam * cv.throughput * n.radiance * geom * mis;
  if (camTerminated && lightTerminated) r.undo_secondary_rescale();
  out.contrib = r;
  return true;
}

// The light-vertex context that the connection to light vertex j inherits from vertices 0 .. j-1
// (LightVertexContext, advanced as the connection loop of traceCameraRays walks the subpath).
struct LightPrefix {
  LightVertexCtx lv;
  float deferred;
  YPT_HD void begin() {
    lv.riSum = 0.f;
    lv.prevFlags = 0;
    lv.originFlags = 0;
    lv.prevPos = vec3(0.f);
    lv.prevNrm = vec3(0.f);
    lv.prevFwdMIS = 0.f;
    deferred = 0.f;
  }
  // past light vertex n (index j): the context of vertex j + 1
  YPT_HD void advance(const LightNode& n, uint32_t j) {
    lv.riSum = deferred;
    deferred = accumulate_ri(deferred, n.pdfBwd, n.pdfFwd, (n.flags & LN_SUBSURFACE_EXIT) == 0);
    lv.prevNrm = n.nrm;
    lv.prevPos = n.pos;
    lv.prevFwdMIS = n.pdfFwd;
    lv.prevFlags = n.flags;
    if (j == 0) lv.originFlags = n.flags;
  }
};

// One step of the connection loop: camera vertex cv (camera vertex index k, surface ctx, path state
// rs) to light vertex n = j of the subpath, whose inherited context is lv. The direction cv.wi is set
// here on the copy. False when the strategy does not contribute or the connection carries nothing.
template <class Surfaces>
YPT_HD bool connect_camera_to_node(const SceneView& S, const FrameConstants& fc, uint32_t k, CameraVertexCtx cv,
                                   const SurfaceContext& ctx, const BdptState& rs, const Lambdas& L, const LightNode& n, uint32_t j,
                                   const LightVertexCtx& lv, Surfaces& surfs, PendingConnection& pc) {
  if ((n.flags & LN_DELTA) != 0 || !should_contribute(fc, k + 1u, j + 1u)) return false;
  if (j == 0 && infinite_node(n.flags)) {
    cv.wi = normalize(-infinite_node_direction(S, n));
  } else {
    vec3 toNode = n.pos - cv.pos;
    float dist = length(toNode);
    cv.wi = normalize(toNode / dist);
  }
  const LightNodeSurface* ls = j > 0 ? surfs.get(S, n, j, lv.prevPos, L) : nullptr;
  return connection_prepare(S, fc, k, cv, ctx, n, j, ls, lv, rs, L, pc);
}

// A camera vertex whose connections a separate launch makes, one thread per light vertex (the
// bidirectional camera wavefront): the vertex context, its surface and path state as the connection
// loop would read them, and how many light vertices to connect to (0: none).
struct DeferredConnections {
  CameraVertexCtx cv;
  SurfaceContext ctx;
  BdptState rs;
  uint32_t k;
  uint32_t count;
};

// where camera_vertex's connections go: made in its own loop (ConnectInLoop) or left, with the
// vertex context, to a separate launch (ConnectLater)
struct ConnectInLoop {
  static constexpr bool later = false;
  DeferredConnections* d = nullptr;
};
struct ConnectLater {
  static constexpr bool later = true;
  DeferredConnections* d;
};

// ---------------------------------------------------------------- camera subpath
// traceCameraRays (BDPTTraceCameraPath.comp:40-432) as a resumable state machine: camera_begin sets
// up texel (px, py), camera_vertex consumes the traced ray's hit (or miss) and does one iteration of
// the vertex loop, handing every connection with its visibility ray to `sink` in light-vertex order.
struct CameraPath {
  BdptState rs;
  Spectrum throughput, total;
  vec3 rayOrigin, rayDir;
  vec3 prevPos, prevNrm;
  float pdfRatioSum, prevFwdMIS;
  float pdfForwardCamera, deferredCam;
  float lambda;
  uint32_t dim, seed;
  uint32_t k;      // camera vertex index of the ray in flight
  bool alive;
  bool pendingExit;    // the next vertex is the exit point of the subsurface event just taken (no ray)
  bool prevIsExit;     // the previous vertex was a subsurface exit
  SurfacePoint exitP;
};

// extraDim: the workgroup's first thread drew the fallback wavelength from its own sampler when
// there were no light subpaths (BDPTTraceCameraPath.comp:67-74)
YPT_HD void camera_begin(const FrameConstants& fc, const uint32_t* point, uint32_t px, uint32_t py, float lambda, bool extraDim,
                         CameraPath& c) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0];
  RandomSampler rand;
  rand.init(point, camera_dim_offset(fc), py * W + px);
  float lens[2];
  rand.r2(lens);
  if (extraDim) rand.dim += 1u;
  float jx, jy;
  rand.texel_offset(jx, jy);
  camera_ray(fc, px, py, jx, jy, lens[0], lens[1], c.rayOrigin, c.rayDir);
  c.rs.init();
  c.throughput.set(1.f);
  c.total.set(0.f);
  c.prevPos = vec3(0.f);
  c.prevNrm = vec3(0.f);
  c.pdfRatioSum = 0.f;
  c.prevFwdMIS = 0.f;
  c.pdfForwardCamera = camera_dir_pdf(fc, c.rayDir);
  c.deferredCam = 0.f;
  c.lambda = lambda;
  c.dim = rand.dim;
  c.seed = rand.seed;
  c.k = 0;
  c.alive = fc.maxRayDepth > 0;
  c.pendingExit = false;
  c.prevIsExit = false;
}

// A vertex is the traced ray's hit or miss, or (c.pendingExit) the exit point of the subsurface
// event taken at the previous vertex; `probe` traces the subsurface probe rays.
template <class Sink, class Surfaces, class Probe, class Connect = ConnectInLoop>
YPT_HD void camera_vertex(const SceneView& S, const FrameConstants& fc, const uint32_t* point, CameraPath& c, bool hit,
                          const Hit& h, const LightNodePacked* nodes, uint32_t nodeCount, Surfaces& surfs, Sink& sink,
                          FirstHitFeatures* fh, Probe& probe, Connect connect = Connect()) {
  if constexpr (Connect::later) connect.d->count = 0;
  RandomSampler rand;
  rand.point = point;
  rand.dim = c.dim;
  rand.seed = c.seed;
  Lambdas L = spectral_lambdas(c.lambda);
  const uint32_t k = c.k;
  float rs0[4];
  rand.r4(rs0);
  c.alive = false;
  CameraVertexCtx cv;
  cv.throughput = c.throughput;
  cv.prevPos = c.prevPos;
  cv.prevNrm = c.prevNrm;
  cv.pdfRatioSum = c.pdfRatioSum;
  cv.prevFwdMIS = c.prevFwdMIS;
  const bool isExit = c.pendingExit;
  c.pendingExit = false;
  cv.isExit = isExit;
  cv.prevIsExit = c.prevIsExit;
  if (!hit && !isExit) {
    cv.curFwdMIS = c.pdfForwardCamera;
    float w = mis_camera_misses(fc, k, cv, c.rayDir);
    if (should_contribute(fc, k + 1u, 0u))
      c.total = c.total + cv.throughput * rgb_unbounded_spectrum(S, sky_color(S, fc, c.rayDir), L) * w;
    c.dim = rand.dim;
    return;
  }
  if (k >= fc.regularizeAfterVertices) c.rs.flags |= RS_REGULARIZE;
  BdptVertex v;
  const MaterialGPU* m = nullptr;
  if (isExit) bdpt_surface_exit(S, c.rs, c.exitP, L, v, &m);
  else bdpt_surface(S, c.rs, h, c.rayOrigin, c.rayDir, L, v, &m);
  if (k == 0 && fh) {
    SurfaceTexels tx;
    sample_surface_texels(S, *m, h.tri, h.b1, h.b2, tx);
    first_hit_features(fc, *m, v.ctx.def, v.pos, *fh, &tx);
  }
  cv.pos = v.pos;
  cv.frontFace = v.frontFace;
  cv.wo = v.wo;
  cv.matId = v.matId;
  cv.normalWS = v.normalWS;
  // a subsurface exit only has the exit lobe, and its edge density is one of area already
  bool delta = !isExit && is_delta_distribution(v.ctx.pre.ax, v.ctx.pre.ay, v.ctx.def);
  if (isExit) {
    cv.curFwdMIS = c.pdfForwardCamera;
  } else {
    cv.curFwdMIS = c.pdfForwardCamera < 0.f ? -1.f : c.pdfForwardCamera * solid_angle_to_area(v.pos - c.rayOrigin, v.normalWS);
    Spectrum ab = c.rs.get_absorption();
    if (!ab.all_equal(0.f)) cv.throughput = cv.throughput * transmittance(h.t, ab);
  }
  const SurfaceDef& sd = v.ctx.def;
  if (!isExit && !sd.emissive.all_equal(0.f)) {  // emission at an exit was counted at the entry
    float w = mis_camera_hits_light(S, k, h.tri, cv);
    if (should_contribute(fc, k + 1u, 0u) && (surface_two_sided(sd.flags) || dot(sd.geometryNormal, v.wo) > 0.f)) {
      float profile = emission_profile(dot(v.normalWS, -c.rayDir), sd.emissionFocus);
      c.total = c.total + cv.throughput * sd.emissive * w * profile;
    }
    c.throughput = cv.throughput;
    c.dim = rand.dim;
    return;
  }
  if constexpr (Connect::later) {  // connections made by a separate launch from what the loop would read
    if (!delta) {
      connect.d->cv = cv;
      connect.d->ctx = v.ctx;
      connect.d->rs = c.rs;
      connect.d->k = k;
      connect.d->count = nodeCount;
    }
  } else if (!delta) {
    LightPrefix pre;
    pre.begin();
    for (uint32_t j = 0; j < nodeCount; ++j) {
      LightNode n = unpack_node(nodes[j]);
      PendingConnection pc;
      if (connect_camera_to_node(S, fc, k, cv, v.ctx, c.rs, L, n, j, pre.lv, surfs, pc)) sink(c, pc);
      pre.advance(n, j);
    }
  }
  c.throughput = cv.throughput;
  ScatterEvent se;
  sample_scattering_event(S, v, isExit, rs0, probe, se);
  if (se.type == EV_NONE) { c.dim = rand.dim; return; }
  if (se.isDiffuse) c.rs.flags |= RS_REGULARIZE;
  ScatterEval ev;
  bool bwdIsArea = false;
  if (se.type == EV_SSS_ENTRY) {
    ev = eval_subsurface_entry(S, v, se.exit);
  } else if (se.type == EV_SSS_EXIT_LOBE) {
    ev = eval_subsurface_exit(S, v.ctx, c.rs, v.matId, v.wo, se.wi, v.frontFace, false, v.pos, v.normalWS, c.prevPos, c.prevNrm);
    bwdIsArea = true;
  } else {
    ev = eval_and_pdfs(S, v.ctx, c.rs, v.matId, v.wo, se.wi, v.frontFace, false, se.isDelta);
  }
  Spectrum sampleWeight = ev.weight * safe_div(1.f, ev.pdfFwd);
  if (ev.dispersed) c.throughput.terminate_secondary();
  if (c.throughput.russian_roulette(rand.r1())) { c.dim = rand.dim; return; }
  if (k > 0) {
    float pb = se.isDelta ? -1.f : (bwdIsArea ? ev.pdfBwd : ev.pdfBwd * solid_angle_to_area(c.rayOrigin - v.pos, c.prevNrm));
    c.deferredCam = accumulate_ri(c.deferredCam, pb, c.prevFwdMIS, !c.prevIsExit);
    c.pdfRatioSum = c.deferredCam;
  }
  c.prevNrm = v.normalWS;
  c.prevPos = v.pos;
  c.prevFwdMIS = cv.curFwdMIS;
  c.prevIsExit = isExit;
  c.pdfForwardCamera = se.isDelta ? -1.f : ev.pdfFwd;
  c.throughput = c.throughput * sampleWeight;
  if (se.type == EV_SSS_ENTRY) {
    c.pendingExit = true;
    c.exitP = se.exit.p;
  } else {
    c.rayDir = normalize(se.wi);
    c.rayOrigin = spawn_ray_origin(v.pos, v.normalWS, c.rayDir);
  }
  c.dim = rand.dim;
  c.k = k + 1;
  c.alive = c.k < fc.maxRayDepth;
}

template <class Tracer>
struct TraceNowSink {  // tests visibility at once and adds visible contributions in order
  Tracer* tr;
  YPT_HD void operator()(CameraPath& c, const PendingConnection& pc) {
    bool visible = pc.tri == NO_HIT ? tr->unoccluded(pc.origin, pc.dir, RAY_MIN_T, pc.tmax)
                                       : tr->hits_target(pc.origin, pc.dir, RAY_MIN_T, pc.tmax, pc.tri);
    if (visible) c.total = c.total + pc.contrib;
  }
};

// the whole camera subpath of texel (px, py) with inline tracing
template <class Tracer, class Surfaces>
YPT_HD Spectrum trace_camera_path(const SceneView& S, const FrameConstants& fc, const uint32_t* point, uint32_t px, uint32_t py,
                                  float lambda, const LightNodePacked* nodes, uint32_t nodeCount, Surfaces& surfs, bool extraDim,
                                  Tracer& tr, FirstHitFeatures* fh) {
  CameraPath c;
  camera_begin(fc, point, px, py, lambda, extraDim, c);
  TraceNowSink<Tracer> sink;
  sink.tr = &tr;
  while (c.alive) {
    Hit h;
    bool hit = !c.pendingExit && tr.closest(c.rayOrigin, c.rayDir, RAY_MIN_T, __builtin_inff(), h);
    camera_vertex(S, fc, point, c, hit, h, nodes, nodeCount, surfs, sink, c.k == 0 ? fh : nullptr, tr);
  }
  return c.total;
}

// light subpath the camera workgroup `group` connects to (BDPTTraceCameraPath.comp:75-83)
YPT_HD uint32_t pick_light_path(const uint32_t* point, uint32_t group, uint32_t headerCount) {
  RandomSampler r;
  r.init(point, 0u, group);
  float u = r.r1();
  return (uint32_t)fmin_(floorf(u * (float)headerCount), (float)(headerCount - 1u));
}

// wavelength of a workgroup with no light subpaths: the first thread's sampler after the lens sample
YPT_HD float fallback_wavelength(const FrameConstants& fc, const uint32_t* point, uint32_t px, uint32_t py) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0];
  RandomSampler rand;
  rand.init(point, camera_dim_offset(fc), py * W + px);
  rand.dim += 2u;
  return hero_wavelength(rand.r1());
}

// ---------------------------------------------------------------- light tracing onto the lens
// projectThroughLensToUV (raytraceCommonResources.hlsl:139-156)
YPT_HD bool project_to_uv(const FrameConstants& fc, vec3 lensPosVS, vec3 posWS, float& u, float& v) {
  vec3 posVS = xform_point(fc.worldToView, posWS);
  vec3 lensToPos = posVS - lensPosVS;
  float forwardDist = lensToPos.z * fc.lensParams[2];
  if (forwardDist <= 0.f) { u = 0.f; v = 0.f; return false; }
  vec3 focus = lensPosVS + lensToPos * (fc.lensParams[1] / forwardDist);
  float in[4] = {focus.x, focus.y, focus.z, 1.f}, o[4];
  mul4(fc.viewToUV, in, o);
  u = o[0] / o[3];
  v = o[1] / o[3];
  return u >= 0.f && v >= 0.f && u <= 1.f && v <= 1.f;
}

struct SplatRecord {
  uint32_t texel;  // npix when the vertex adds nothing
  vec3 xyz;
};

// calculatePureLightPaths (BDPTCalculatePureLightPaths.comp:107-206), in pieces: the lens point of a
// light subpath (splat_lens), the MIS state that vertex i inherits from vertices 0 .. i-1
// (SplatPrefix, advanced one vertex at a time), and the splat of one vertex (splat_vertex).
// splat_light_path runs them in vertex order; the per-vertex launch of the GPU computes each vertex's
// prefix with the same updates, so both give the same bits.
struct SplatLens {
  vec3 posVS, posWS;
};

YPT_HD void splat_lens(const FrameConstants& fc, const uint32_t* point, uint32_t header, SplatLens& out) {
  RandomSampler lr;
  lr.init(point, pure_lens_dim_offset(fc), header);
  float u2[2];
  lr.r2(u2);
  float dx, dy;
  sample_concentric_disk(u2[0], u2[1], dx, dy);
  out.posVS = vec3(dx * fc.lensParams[0], dy * fc.lensParams[0], 0.f);
  out.posWS = xform_point(fc.viewToWorld, out.posVS);
}

struct SplatPrefix {
  float riSum, deferred;
  vec3 prevPos, prevNrm;
  float prevFwd;
  uint32_t prevFlags;
  YPT_HD void begin(const LightNode& n0) {
    riSum = 0.f;
    deferred = accumulate_ri(0.f, n0.pdfBwd, n0.pdfFwd);
    prevPos = n0.pos;
    prevNrm = n0.nrm;
    prevFwd = n0.pdfFwd;
    prevFlags = n0.flags;
  }
  // past vertex n (index >= 1): the state its successor inherits
  YPT_HD void advance(const LightNode& n) {
    riSum = deferred;
    deferred = accumulate_ri(deferred, n.pdfBwd, n.pdfFwd, (n.flags & LN_SUBSURFACE_EXIT) == 0);
    prevPos = n.pos;
    prevNrm = n.nrm;
    prevFwd = n.pdfFwd;
    prevFlags = n.flags;
  }
};

// the splat of light vertex i >= 1 (rec.texel = npix when it adds nothing)
template <class Tracer, class Surfaces>
YPT_HD void splat_vertex(const SceneView& S, const FrameConstants& fc, uint32_t headerCount, const Lambdas& L, const SplatLens& lens,
                         const LightNode& n0, const LightNode& n, uint32_t i, const SplatPrefix& pf, Surfaces& surfs, Tracer& tr,
                         SplatRecord& rec) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1];
  rec.texel = W * H;
  float u, v;
  if (!(project_to_uv(fc, lens.posVS, n.pos, u, v) && (n.flags & LN_DELTA) == 0 && should_contribute(fc, 0u, i + 1u))) return;
  ScatterEval ev;
  float unused;
  if (!connect_light_node(S, fc, n, i, surfs.get(S, n, i, pf.prevPos, L), lens.posWS, false, L, ev, unused, pf.prevPos, pf.prevNrm) ||
      ev.weight.all_equal(0.f))
    return;
  // checkVisibility (:70-101): from the vertex to the lens point, t in [0.01, distance]
  vec3 fromTo = lens.posWS - n.pos;
  float len = sqrtf(dot(fromTo, fromTo));
  bool visible = len < 0.0001f || tr.unoccluded(n.pos, fromTo / len, 0.01f, len);
  if (!visible) return;
  bool prevEnv = i == 1 && (n0.flags & LN_ENV) != 0;
  bool prevPunctual = i == 1 && (n0.flags & LN_PUNCTUAL) != 0;
  bool nodeExit = (n.flags & LN_SUBSURFACE_EXIT) != 0;  // its backward density is one of area already
  float pbPrev = (prevEnv || nodeExit) ? ev.pdfBwd : ev.pdfBwd * solid_angle_to_area(pf.prevPos - n.pos, pf.prevNrm);
  if (prevPunctual) pbPrev = 0.f;
  vec3 camToNode = n.pos - lens.posWS;
  float pdfCam = camera_dir_pdf(fc, camToNode);
  float ri = accumulate_ri(pf.riSum, pbPrev, pf.prevFwd, (pf.prevFlags & LN_SUBSURFACE_EXIT) == 0);
  ri = accumulate_ri(ri, pdfCam * solid_angle_to_area(camToNode, n.nrm), n.pdfFwd, !nodeExit);
  float mis = 1.f / (1.f + ri);
  float geom = 1.f / fmax_(0.0001f, dot(camToNode, camToNode));
  float importance = safe_div(pdfCam, (float)headerCount);
  Spectrum c = n.radiance * ev.weight * mis * geom * importance;
  float su = fmin_(fmax_(u, 0.f), 1.f), sv = fmin_(fmax_(v, 0.f), 1.f);
  uint32_t tx = (uint32_t)floorf(fmin_(su * (float)W, (float)(W - 1u)));
  uint32_t ty = (uint32_t)floorf(fmin_(sv * (float)H, (float)(H - 1u)));
  rec.texel = ty * W + tx;
  rec.xyz = spectrum_to_xyz(c, L, S.spec);
}

// the light subpath in header slot `header`: one record per vertex 1 .. count-1 into rec[0 .. count-1)
template <class Tracer, class Surfaces>
YPT_HD void splat_light_path(const SceneView& S, const FrameConstants& fc, const uint32_t* point, uint32_t header,
                             uint32_t headerCount, float lambda, const LightNodePacked* nodes, uint32_t count, Surfaces& surfs,
                             Tracer& tr, SplatRecord* rec) {
  uint32_t npix = (uint32_t)fc.targetTexDimensions[0] * (uint32_t)fc.targetTexDimensions[1];
  for (uint32_t i = 1; i < count; ++i) rec[i - 1].texel = npix;
  if (count <= 1) return;
  Lambdas L = spectral_lambdas(lambda);
  SplatLens lens;
  splat_lens(fc, point, header, lens);
  LightNode n0 = unpack_node(nodes[0]);
  SplatPrefix pf;
  pf.begin(n0);
  for (uint32_t i = 1; i < count; ++i) {
    LightNode n = unpack_node(nodes[i]);
    splat_vertex(S, fc, headerCount, L, lens, n0, n, i, pf, surfs, tr, rec[i - 1]);
    pf.advance(n);
  }
}

// Texel permutation of the camera dispatch (BidirectionalPathIntegratorSubStage.cpp:549-591): every
// 8x8 workgroup covers 64 texels spread over the image. Each position is swapped with a random one,
// drawn from a seeded SplitMix64 stream here.
inline void texel_permutation(uint32_t W, uint32_t H, uint64_t seed, uint32_t* out) {
  uint32_t n = W * H;
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t x = 0; x < W; ++x) out[y * W + x] = (y << 16) | (x & 0xFFFFu);
  uint64_t s = seed;
  for (uint32_t i = 0; i < n; ++i) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    uint32_t j = (uint32_t)(z % n);
    uint32_t t = out[i]; out[i] = out[j]; out[j] = t;
  }
}

}  // namespace ypt
