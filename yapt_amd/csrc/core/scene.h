// Scene and per-frame data as the kernels and the host build see it.
//
// Geometry is a world-space triangle soup (render-object transforms baked at upload, which makes
// object space equal world space in the shading code of rayHitDefault.hlsl:464-585). An "instance"
// is one (render object, submesh) pair as in the reference's TLAS (AccelerationStructureHelper.cpp),
// owning a contiguous triangle range, one material and optionally one mesh light.
#pragma once
#include "material.h"

namespace ypt {

struct InstanceGPU {
  uint32_t material;
  uint32_t firstTri;
  uint32_t triCount;
  float lightSelectionPdf;  // g_instanceLightSelectionPDFs: 0 for instances that are not lights
};

// light types (lightTypes.h:5-8)
static constexpr uint32_t LIGHT_TYPE_MESH = 0;
static constexpr uint32_t LIGHT_TYPE_POINT = 1;
static constexpr uint32_t LIGHT_TYPE_SPOT = 2;
static constexpr uint32_t LIGHT_TYPE_DIRECTIONAL = 3;

// LightEntryGPU (materialsCommonResources.hlsl:51-71): mesh lights first, then the punctual lights
// (RaytraceCommonResources.cpp:452-545).
struct LightGPU {
  uint32_t instance;     // mesh lights: the emitting instance and its triangle range
  uint32_t material;
  uint32_t firstTri;
  uint32_t triCount;
  float selectionPdf;    // alias table over all lights (RaytraceCommonResources.cpp:325-368)
  float aliasThreshold;
  uint32_t aliasIndex;
  uint32_t type;         // LIGHT_TYPE_*
  float positionCosInner[4];   // punctual lights
  float directionCosOuter[4];  // emission direction (normalized), cosine of the outer spot angle
  float intensity[4];          // RGB color * intensity
  float punctualSelectionPdf;  // alias table over the punctual lights only
  float punctualAliasThreshold;
  uint32_t punctualAliasIndex;
  uint32_t pad;
};

static constexpr uint32_t INTEGRATOR_BPT = 0;
static constexpr uint32_t INTEGRATOR_NEE = 1;
static constexpr uint32_t INTEGRATOR_BDPT = 2;

static constexpr uint32_t FRAME_FLAG_DISABLE_JITTER = 1u << 1;  // CONSTANTS_FLAG_DISABLE_TEXEL_JITTER
static constexpr uint32_t FRAME_FLAG_WHITE_FURNACE = 1u << 8;   // WHITE_FURNACE_TEST as a runtime switch

// RaytraceConstantData (raytraceCommonResources.hlsl:8-31) plus the integrator parameters.
struct FrameConstants {
  mat4 uvToView, viewToUV, worldToView, viewToWorld;
  float worldBoundsMin[4], worldBoundsMax[4];
  float cameraPosition[4];
  float targetTexDimensions[4];  // W, H, 1/W, 1/H
  float rayUVOffset[2];
  uint32_t maxRayDepth;
  uint32_t flags;
  float envIntensityScale;
  uint32_t lightCount;
  uint32_t regularizeAfterVertices;
  uint32_t integrator;
  float lensParams[4];  // aperture radius, focus distance, view-space forward z sign, unused
  // bidirectional integrator (BidirectionalPathTraceConstants, BDPTCommon.hlsl:30-38)
  uint32_t lightPathVertices;  // maxVerticesPerLightPath
  uint32_t lightPathsPerDim[2];
  uint32_t punctualLightCount;  // the last punctualLightCount entries of the light table
};

struct SceneView {
  const float* pos;          // 3 floats per vertex
  const float* nrm;          // 3 floats per vertex
  const float* tan;          // 4 floats per vertex (xyz, handedness), or null
  const uint32_t* idx;       // 3 per triangle
  const uint32_t* triInst;   // instance of each triangle
  const InstanceGPU* inst;
  const MaterialGPU* mats;
  const LightGPU* lights;
  const float* rgb2spec;     // optional 64^3 x 4 RGB->coefficient table (environment colours)
  uint32_t ntris;
  uint32_t ninst;
  SpectralTables spec;
  MSLuts luts;
};

YPT_HD vec3 load3(const float* p, uint32_t i) { return vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

YPT_HD void tri_vertices(const SceneView& S, uint32_t tri, vec3& p1, vec3& p2, vec3& p3) {
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  p1 = load3(S.pos, i0); p2 = load3(S.pos, i1); p3 = load3(S.pos, i2);
}

}  // namespace ypt
