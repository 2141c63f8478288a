// Scene and per-frame data as the kernels and the host build see it.
//
// Geometry is a world-space triangle soup (render-object transforms baked at upload, which makes
// object space equal world space in the shading code of rayHitDefault.hlsl:464-585). An "instance"
// is one (render object, submesh) pair as in the reference's TLAS (AccelerationStructureHelper.cpp),
// owning a contiguous triangle range, one material and optionally one mesh light.
#pragma once
#include "material.h"
#include "ymath.h"

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
  // strategies that contribute: camera vertices in [nodeRanges[2], nodeRanges[3]] and light vertices in
  // [nodeRanges[0], nodeRanges[1]] (lightAndCameraPathConstraints, BDPTCommon.hlsl:233-234)
  uint32_t nodeRanges[4];
};

// one RGBA8 texture in the texel buffer (ypt_texture)
struct TextureGPU {
  uint32_t offset;  // first texel
  uint32_t width, height;
  uint32_t srgb;    // decode sRGB before filtering
};

// Alpha testing (hitShadersCommon.hlsl:162-183): a candidate hit on a triangle whose material has
// an alpha cutoff, an albedo texture and UVs is taken only where the albedo texture's alpha at the
// hit is at least the cutoff, for closest hits and shadow queries alike. Each such triangle has a
// record with its vertex UVs, so that traversal samples the alpha without the mesh buffers; the UV
// interpolation is the one surface_uv does and the alpha the one sample_texture filters.
struct alignas(16) AlphaTri {
  float uv[6];     // vertex UVs
  float scale[2];  // albedo texture UV scale
  uint32_t tex;    // albedo texture
  float cutoff;
  uint32_t pad[2];
};

struct AlphaView {
  const AlphaTri* tris;  // null: the scene has no alpha-tested triangles
  const TextureGPU* textures;
  const uint32_t* texels;
};

// Environment texture, the World.Skycube renderer variable (RVARNAME_SKYBOX, CRenderer.h:18) as
// RaytraceCommonResources.cpp:285-313 classifies it: none, a cube map or a long-lat image
// (ENVIRONMENT_TYPE_*, RaytraceCommonResources.h:17-19).
static constexpr uint32_t ENV_NONE = 0, ENV_CUBE = 1, ENV_LONGLAT = 2;

struct EnvView {
  const float* texels;  // linear RGBA, rows top to bottom; a cube map holds its faces +X, -X, +Y, -Y, +Z, -Z
  uint32_t type;        // ENV_*
  uint32_t width, height;  // of the long-lat image or of one cube face
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
  const float* uv;           // 2 floats per vertex, or null: UV (0.5, 0.5)
  const TextureGPU* textures;
  const uint32_t* texels;    // RGBA8 per texel
  const float* srgbLut;      // 256 entries: sRGB-encoded byte -> linear
  const uint32_t* triAlpha;  // per triangle: 0 opaque, else 1 + its alpha-test record; null: none
  AlphaView alpha;
  uint32_t ntris;
  uint32_t ninst;
  SpectralTables spec;
  MSLuts luts;
  EnvView env;
};

static constexpr float RAY_MIN_T = 5e-4f;  // DEFAULT_RAY_MIN_T (raytraceCommon.hlsl:22-23)

struct Hit {
  float t, b1, b2;  // b1, b2: weights of the second and third vertex (DXR barycentrics)
  uint32_t tri;     // NO_HIT on a miss
};

YPT_HD vec3 load3(const float* p, uint32_t i) { return vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

YPT_HD void tri_vertices(const SceneView& S, uint32_t tri, vec3& p1, vec3& p2, vec3& p3) {
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  p1 = load3(S.pos, i0); p2 = load3(S.pos, i1); p3 = load3(S.pos, i2);
}

// Texture2D.SampleLevel(g_colorSampler, uv, 0): bilinear between texel centres, wrap addressing,
// sRGB textures decoded per texel before the filter as an _SRGB format is
// (the decoding table exists whenever the scene has textures; it is read for every texel and the
// value kept for sRGB textures, so that no branch separates the table reads)
YPT_HD void texel_decode(const SceneView& S, const TextureGPU& t, uint32_t v, float out[4]) {
  for (int k = 0; k < 3; ++k) {
    uint32_t b = (v >> (8 * k)) & 0xFFu;
    float lin = S.srgbLut[b];
    out[k] = t.srgb ? lin : (float)b / 255.f;
  }
  out[3] = (float)(v >> 24) / 255.f;
}

YPT_HD void sample_texture(const SceneView& S, uint32_t index, float u, float v, float out[4]) {
  const TextureGPU& t = S.textures[index];
  float uf = u - floorf(u), vf = v - floorf(v);  // wrap first: keeps the texel index in range
  float x = uf * (float)t.width - 0.5f, y = vf * (float)t.height - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  uint32_t x0 = fx < 0.f ? t.width - 1u : (uint32_t)fx, y0 = fy < 0.f ? t.height - 1u : (uint32_t)fy;
  if (x0 >= t.width) x0 = t.width - 1u;
  if (y0 >= t.height) y0 = t.height - 1u;
  uint32_t x1 = x0 + 1u == t.width ? 0u : x0 + 1u, y1 = y0 + 1u == t.height ? 0u : y0 + 1u;
  // the four texel words first, so their loads are in flight together, then their decoding
  const uint32_t* row0 = S.texels + t.offset + y0 * t.width;
  const uint32_t* row1 = S.texels + t.offset + y1 * t.width;
  uint32_t w00 = row0[x0], w10 = row0[x1], w01 = row1[x0], w11 = row1[x1];
  float a[4], b[4], c[4], d[4];
  texel_decode(S, t, w00, a);
  texel_decode(S, t, w10, b);
  texel_decode(S, t, w01, c);
  texel_decode(S, t, w11, d);
  for (int k = 0; k < 4; ++k) {
    float top = a[k] + ax * (b[k] - a[k]);
    float bot = c[k] + ax * (d[k] - c[k]);
    out[k] = top + ay * (bot - top);
  }
}

// Bilinear filter of one environment image (a long-lat image or a cube face) at (u, v), between texel
// centres: wrap addressing for the long-lat image (the LINEAR_REPEAT sampler, also across the poles, as
// the sampler does), clamp to the face edge for a cube face. Hardware cube sampling filters across face
// edges; here a lookup within half a texel of an edge repeats the edge texels instead.
YPT_HD vec3 env_bilinear(const EnvView& e, uint32_t face, float u, float v, bool wrap) {
  const uint32_t W = e.width, H = e.height;
  if (wrap) {
    u = u - floorf(u);
    v = v - floorf(v);
  }
  if (!(u == u)) u = 0.f;  // a NaN or infinite coordinate (a degenerate direction) reads a corner texel
  if (!(v == v)) v = 0.f;
  float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  uint32_t x0, x1, y0, y1;
  if (wrap) {
    x0 = fx < 0.f ? W - 1u : (uint32_t)fx;
    y0 = fy < 0.f ? H - 1u : (uint32_t)fy;
    if (!(x0 < W)) x0 = W - 1u;
    if (!(y0 < H)) y0 = H - 1u;
    x1 = x0 + 1u == W ? 0u : x0 + 1u;
    y1 = y0 + 1u == H ? 0u : y0 + 1u;
  } else {
    x0 = fx < 0.f ? 0u : (fx >= (float)(W - 1u) ? W - 1u : (uint32_t)fx);
    y0 = fy < 0.f ? 0u : (fy >= (float)(H - 1u) ? H - 1u : (uint32_t)fy);
    x1 = fx + 1.f >= (float)(W - 1u) ? W - 1u : (fx + 1.f < 0.f ? 0u : (uint32_t)(fx + 1.f));
    y1 = fy + 1.f >= (float)(H - 1u) ? H - 1u : (fy + 1.f < 0.f ? 0u : (uint32_t)(fy + 1.f));
  }
  const float* base = e.texels + (size_t)face * W * H * 4u;
  const float* a = base + 4u * ((size_t)y0 * W + x0);
  const float* b = base + 4u * ((size_t)y0 * W + x1);
  const float* c = base + 4u * ((size_t)y1 * W + x0);
  const float* d = base + 4u * ((size_t)y1 * W + x1);
  float out[3];
  for (int k = 0; k < 3; ++k) {
    float top = a[k] + ax * (b[k] - a[k]);
    float bot = c[k] + ax * (d[k] - c[k]);
    out[k] = top + ay * (bot - top);
  }
  return vec3(out[0], out[1], out[2]);
}

// Cube-map face and face coordinates of a direction: the major axis picks the face (z before y before x
// on ties, as the hardware's cube-id rule), then (s, t) = ((sc / |ma| + 1) / 2, (tc / |ma| + 1) / 2) with
// the face table shared by D3D and Vulkan (+X: -z, -y; -X: z, -y; +Y: x, z; -Y: x, -z; +Z: x, -y;
// -Z: -x, -y).
YPT_HD void env_cube_coords(vec3 d, uint32_t& face, float& s, float& t) {
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  float ma, sc, tc;
  if (az >= ax && az >= ay) {
    bool pos = !(d.z < 0.f);
    face = pos ? 4u : 5u;
    ma = az;
    sc = pos ? d.x : -d.x;
    tc = -d.y;
  } else if (ay >= ax) {
    bool pos = !(d.y < 0.f);
    face = pos ? 2u : 3u;
    ma = ay;
    sc = d.x;
    tc = pos ? d.z : -d.z;
  } else {
    bool pos = !(d.x < 0.f);
    face = pos ? 0u : 1u;
    ma = ax;
    sc = pos ? -d.z : d.z;
    tc = -d.y;
  }
  s = 0.5f * (sc / ma + 1.f);
  t = 0.5f * (tc / ma + 1.f);
}

// getSkyBoxColor (materialsCommonResources.hlsl:181-223), RGB: the long-lat image at
// (atan2(z, x) / 2pi + 1/2, acos(y) / pi) of the normalised direction (cartesianToSpherical,
// commonMath.hlsl:144-151), the cube map along the direction, or |dir| without a texture
YPT_HD vec3 env_color(const SceneView& S, vec3 dir) {
  const EnvView& e = S.env;
  if (e.type == ENV_LONGLAT) {
    vec3 n = normalize(dir);
    float theta = yatan2(n.z, n.x);
    float phi = yacos(clampf(n.y, -1.f, 1.f));
    return env_bilinear(e, 0u, theta * INVPI * 0.5f + 0.5f, phi * INVPI, true);
  }
  if (e.type == ENV_CUBE) {
    uint32_t face;
    float s, t;
    env_cube_coords(dir, face, s, t);
    return env_bilinear(e, face, s, t, false);
  }
  return vabs(dir);
}

// fetchMeshUV (materialsCommon.hlsl): interpolated UV, (0.5, 0.5) for meshes without UVs
YPT_HD void surface_uv(const SceneView& S, uint32_t tri, float b1, float b2, float& u, float& v) {
  if (!S.uv) { u = 0.5f; v = 0.5f; return; }
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  float b0 = 1.f - b1 - b2;
  u = b0 * S.uv[2 * i0] + b1 * S.uv[2 * i1] + b2 * S.uv[2 * i2];
  v = b0 * S.uv[2 * i0 + 1] + b1 * S.uv[2 * i1 + 1] + b2 * S.uv[2 * i2 + 1];
}

YPT_HD float texel_alpha(const uint32_t* texels, const TextureGPU& t, uint32_t x, uint32_t y) {
  return (float)(texels[t.offset + y * t.width + x] >> 24) / 255.f;
}

// channel 3 of sample_texture, on the same texel coordinates and weights
YPT_HD float sample_alpha(const TextureGPU& t, const uint32_t* texels, float u, float v) {
  float uf = u - floorf(u), vf = v - floorf(v);
  float x = uf * (float)t.width - 0.5f, y = vf * (float)t.height - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float ax = x - fx, ay = y - fy;
  uint32_t x0 = fx < 0.f ? t.width - 1u : (uint32_t)fx, y0 = fy < 0.f ? t.height - 1u : (uint32_t)fy;
  if (x0 >= t.width) x0 = t.width - 1u;
  if (y0 >= t.height) y0 = t.height - 1u;
  uint32_t x1 = x0 + 1u == t.width ? 0u : x0 + 1u, y1 = y0 + 1u == t.height ? 0u : y0 + 1u;
  float a = texel_alpha(texels, t, x0, y0), b = texel_alpha(texels, t, x1, y0);
  float c = texel_alpha(texels, t, x0, y1), d = texel_alpha(texels, t, x1, y1);
  float top = a + ax * (b - a);
  float bot = c + ax * (d - c);
  return top + ay * (bot - top);
}

// record k (the triangle's alpha word minus one) passes at barycentrics (b1, b2)
YPT_HD bool alpha_test_passes(const AlphaView& A, uint32_t k, float b1, float b2) {
  const AlphaTri& r = A.tris[k];
  float b0 = 1.f - b1 - b2;
  float u = b0 * r.uv[0] + b1 * r.uv[2] + b2 * r.uv[4];
  float v = b0 * r.uv[1] + b1 * r.uv[3] + b2 * r.uv[5];
  return sample_alpha(A.textures[r.tex], A.texels, u * r.scale[0], v * r.scale[1]) >= r.cutoff;
}

// a sampled point of an alpha-tested emitter that the alpha test cuts away emits nothing
YPT_HD bool alpha_cut(const SceneView& S, uint32_t tri, float b1, float b2) {
  if (!S.triAlpha) return false;
  uint32_t a = S.triAlpha[tri];
  return a != 0u && !alpha_test_passes(S.alpha, a - 1u, b1, b2);
}

// the bound textures of material m at a surface point (tri, b1, b2); mask 0 when it has none
YPT_HD void sample_surface_texels(const SceneView& S, const MaterialGPU& m, uint32_t tri, float b1, float b2, SurfaceTexels& tx) {
  tx.mask = 0;
  tx.rgb2spec = S.rgb2spec;
  if (m.tex[0] < 0 && m.tex[1] < 0 && m.tex[2] < 0 && m.tex[3] < 0) return;
  float u, v;
  surface_uv(S, tri, b1, b2, u, v);
  for (int k = 0; k < 4; ++k) {
    if (m.tex[k] < 0) continue;
    sample_texture(S, (uint32_t)m.tex[k], u * m.texScale[2 * k], v * m.texScale[2 * k + 1], tx.v[k]);
    tx.mask |= 1u << k;
  }
}

}  // namespace ypt
