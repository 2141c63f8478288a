// CPU checker for the spectral path-tracing core. TEST INFRASTRUCTURE ONLY: tests/, the smoke check
// and the CPU-baseline leg of bench.py load it; the product never does.
//
// A scalar restatement of the reference's per-pixel integrator loop, one path per thread as the
// DXR ray-generation shader runs it (Source/Renderer/shaders/raytracing/rayGenPrimaryRays.hlsl:10-145),
// with the shadow ray of each explicit light connection traced at once (rayHitDefault.hlsl:83-111:
// the committed closest hit must be the sampled light triangle), followed by the film merge of
// mergeNewSamples.comp:173-265. Surface and light arithmetic (layered BSDF, spectra, sampler,
// camera) comes from the shared FP32 core under yapt_amd/csrc/core, compiled here by g++ with
// -ffp-contract=off; what this file owns is independent of the GPU path:
//   * closest hit by brute force over all triangles, or through its own binned-SAH BVH traversed
//     near child first (both give the minimum (t, triangle id) of the shared watertight test);
//   * the path loop and the per-pixel film merge (clamp against the previous pass's film);
//   * Sobol points straight from the generator matrices.
// Parity unpinned against the reference itself (no per-sample vectors exist); pinned by the tests
// listed in the repository's design notes.
#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../include/yapt_pt.h"
#include "../yapt_amd/csrc/core/intersect.h"
#include "../yapt_amd/csrc/core/shade.h"
#include "../yapt_amd/csrc/core/bdpt.h"
#include "../yapt_amd/csrc/runtime/luts.h"
#include "../yapt_amd/csrc/runtime/prepare.h"
#include "../yapt_amd/csrc/runtime/rgb2spec.h"
#include "../yapt_amd/csrc/data/cie_tables.inc"
#include "../yapt_amd/csrc/data/sobol_matrices.inc"

using namespace ypt;

namespace {

std::string g_err;

struct Tables {
  std::vector<float> msl;
  Rgb2Spec r2s;
  std::vector<float> rgb2specFull;
  Tables() {
    msl.resize(MSL_TOTAL);
    generate_ms_luts(msl.data());
    r2s.init(YPT_CIE_X_D, YPT_CIE_Y_D, YPT_CIE_Z_D, YPT_CIE_D65_D, YPT_SRGB_XYZ_TO_RGB_D, YPT_SRGB_RGB_TO_XYZ_D);
  }
};
Tables& tables() {
  static Tables t;
  return t;
}

// Sobol point of a sample index: RandUtility.inl:33-45 for dimensions 0..255
void sobol_point(uint32_t index, uint32_t out[256]) {
  for (uint32_t d = 0; d < 256; ++d) {
    uint32_t v = 0, idx = index;
    for (uint32_t i = d * 32; idx != 0; idx >>= 1, ++i)
      if (idx & 1) v ^= YPT_SOBOL_MATRICES[i];
    out[d] = v;
  }
}

// ------------------------------------------------------------------ closest hit
struct CpuBvhNode {
  float lo[3], hi[3];
  uint32_t left, right;   // inner: child indices
  uint32_t first, count;  // leaf: triangle range in `order` (count > 0)
};

struct Geometry {
  const PreparedScene* S = nullptr;
  bool useBvh = false;
  std::vector<CpuBvhNode> nodes;
  std::vector<uint32_t> order;

  void tri(uint32_t t, vec3& a, vec3& b, vec3& c) const {
    const float* p = S->pos.data();
    const uint32_t* ix = S->idx.data();
    a = load3(p, ix[3 * t]); b = load3(p, ix[3 * t + 1]); c = load3(p, ix[3 * t + 2]);
  }

  void build(const PreparedScene& s, bool bvh) {
    S = &s;
    useBvh = bvh;
    nodes.clear();
    order.clear();
    if (!bvh) return;
    uint32_t n = (uint32_t)s.triInst.size();
    order.resize(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::vector<float> cent(3 * n);
    for (uint32_t t = 0; t < n; ++t) {
      vec3 a, b, c;
      tri(t, a, b, c);
      for (int k = 0; k < 3; ++k) cent[3 * t + k] = (a[k] + b[k] + c[k]) / 3.f;
    }
    struct Item { uint32_t node, first, count; };
    std::vector<Item> st;
    nodes.push_back(CpuBvhNode());
    st.push_back({0, 0, n});
    while (!st.empty()) {
      Item it = st.back();
      st.pop_back();
      CpuBvhNode nd;
      for (int k = 0; k < 3; ++k) { nd.lo[k] = FLT_MAX; nd.hi[k] = -FLT_MAX; }
      for (uint32_t i = it.first; i < it.first + it.count; ++i) {
        vec3 v[3];
        tri(order[i], v[0], v[1], v[2]);
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) { nd.lo[k] = std::min(nd.lo[k], v[j][k]); nd.hi[k] = std::max(nd.hi[k], v[j][k]); }
      }
      nd.left = nd.right = 0;
      nd.first = it.first;
      nd.count = it.count;
      if (it.count > 4) {  // binned SAH split (16 bins per axis over the centroid bounds); median if none pays
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = it.first; i < it.first + it.count; ++i)
          for (int k = 0; k < 3; ++k) { clo[k] = std::min(clo[k], cent[3 * order[i] + k]); chi[k] = std::max(chi[k], cent[3 * order[i] + k]); }
        const int NB = 16;
        int bestAx = -1;
        float bestCut = 0.f, bestCost = FLT_MAX;
        for (int ax = 0; ax < 3; ++ax) {
          float ext = chi[ax] - clo[ax];
          if (!(ext > 0.f)) continue;
          float blo[NB][3], bhi[NB][3];
          uint32_t bcnt[NB] = {0};
          for (int b = 0; b < NB; ++b)
            for (int k = 0; k < 3; ++k) { blo[b][k] = FLT_MAX; bhi[b][k] = -FLT_MAX; }
          for (uint32_t i = it.first; i < it.first + it.count; ++i) {
            int b = std::min(NB - 1, (int)((cent[3 * order[i] + ax] - clo[ax]) / ext * NB));
            vec3 v[3];
            tri(order[i], v[0], v[1], v[2]);
            bcnt[b]++;
            for (int j = 0; j < 3; ++j)
              for (int k = 0; k < 3; ++k) { blo[b][k] = std::min(blo[b][k], v[j][k]); bhi[b][k] = std::max(bhi[b][k], v[j][k]); }
          }
          auto area = [](const float lo[3], const float hi[3]) {
            float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
            return e0 < 0.f ? 0.f : 2.f * (e0 * e1 + e1 * e2 + e2 * e0);
          };
          float rlo[NB][3], rhi[NB][3];
          uint32_t rc[NB];
          float alo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ahi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
          uint32_t acc = 0;
          for (int b = NB - 1; b >= 1; --b) {
            for (int k = 0; k < 3; ++k) { alo[k] = std::min(alo[k], blo[b][k]); ahi[k] = std::max(ahi[k], bhi[b][k]); }
            acc += bcnt[b];
            for (int k = 0; k < 3; ++k) { rlo[b][k] = alo[k]; rhi[b][k] = ahi[k]; }
            rc[b] = acc;
          }
          float llo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
          uint32_t lc = 0;
          for (int b = 0; b < NB - 1; ++b) {
            for (int k = 0; k < 3; ++k) { llo[k] = std::min(llo[k], blo[b][k]); lhi[k] = std::max(lhi[k], bhi[b][k]); }
            lc += bcnt[b];
            if (lc == 0 || rc[b + 1] == 0) continue;
            float cost = area(llo, lhi) * lc + area(rlo[b + 1], rhi[b + 1]) * rc[b + 1];
            if (cost < bestCost) { bestCost = cost; bestAx = ax; bestCut = clo[ax] + ext * (float)(b + 1) / NB; }
          }
        }
        uint32_t mid;
        if (bestAx >= 0) {
          auto left = [&](uint32_t t) {
            return std::min(NB - 1, (int)((cent[3 * t + bestAx] - clo[bestAx]) / (chi[bestAx] - clo[bestAx]) * NB)) <
                   std::min(NB, (int)((bestCut - clo[bestAx]) / (chi[bestAx] - clo[bestAx]) * NB + 0.5f));
          };
          mid = (uint32_t)(std::stable_partition(order.begin() + it.first, order.begin() + it.first + it.count, left) - order.begin());
        } else {
          mid = it.first;
        }
        if (mid == it.first || mid == it.first + it.count) {  // degenerate: median on the widest axis
          int ax = 0;
          for (int k = 1; k < 3; ++k) if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
          mid = it.first + it.count / 2;
          std::nth_element(order.begin() + it.first, order.begin() + mid, order.begin() + it.first + it.count,
                           [&](uint32_t a, uint32_t b) { return cent[3 * a + ax] < cent[3 * b + ax] || (cent[3 * a + ax] == cent[3 * b + ax] && a < b); });
        }
        uint32_t l = (uint32_t)nodes.size();
        nodes.push_back(CpuBvhNode());
        nodes.push_back(CpuBvhNode());
        nd.left = l; nd.right = l + 1; nd.count = 0;
        st.push_back({l, it.first, mid - it.first});
        st.push_back({l + 1, mid, it.first + it.count - mid});
      }
      nodes[it.node] = nd;
    }
  }

  // minimum (t, id) over triangles hit in (tmin, tmax]; with `range` set, over the triangles
  // [lo, lo + range) only (a subsurface probe's instance)
  bool closest(vec3 o, vec3 d, float tmin, float tmax, Hit& h, uint64_t* work, uint32_t lo = 0u,
               uint32_t range = 0xFFFFFFFFu) const {
    RayPre rp = ray_precompute(o, d);
    float best = tmax;
    uint32_t bid = NO_HIT_ID;
    float bb1 = 0.f, bb2 = 0.f;
    const AlphaView av{S->alphaTris.data(), S->textures.data(), S->texels.data()};
    auto test = [&](uint32_t t) {
      if (t - lo >= range) return;
      vec3 a, b, c;
      tri(t, a, b, c);
      float tt, b1, b2;
      if (intersect_triangle(rp, a, b, c, tmin, best, tt, b1, b2) && hit_closer(tt, t, best, bid) &&
          (S->triAlpha[t] == 0u || alpha_test_passes(av, S->triAlpha[t] - 1u, b1, b2))) {
        best = tt; bid = t; bb1 = b1; bb2 = b2;
      }
    };
    if (!useBvh) {
      uint32_t n = (uint32_t)S->triInst.size();
      for (uint32_t t = 0; t < n; ++t) test(t);
      if (work) work[1] += n;
    } else {
      vec3 inv(1.f / d.x, 1.f / d.y, 1.f / d.z);
      // near child first; a popped node is skipped when its entry lies beyond the closest hit so far
      uint32_t stack[128];
      float entry[128];
      int sp = 0;
      stack[sp] = 0;
      entry[sp++] = tmin;
      while (sp) {
        --sp;
        if (entry[sp] > best * SLAB_FAR_SCALE) continue;  // as conservative as the slab test itself
        const CpuBvhNode& nd = nodes[stack[sp]];
        if (work) work[0]++;
        float tn;
        if (!slab_test(vec3(nd.lo[0], nd.lo[1], nd.lo[2]), vec3(nd.hi[0], nd.hi[1], nd.hi[2]), o, inv, tmin, best, tn)) continue;
        if (nd.count) {
          for (uint32_t i = nd.first; i < nd.first + nd.count; ++i) test(order[i]);
          if (work) work[1] += nd.count;
        } else {
          if (sp + 2 > 128) throw std::runtime_error("oracle BVH stack overflow");
          const CpuBvhNode& a = nodes[nd.left];
          const CpuBvhNode& b = nodes[nd.right];
          float ta = FLT_MAX, tb = FLT_MAX;
          bool ha = slab_test(vec3(a.lo[0], a.lo[1], a.lo[2]), vec3(a.hi[0], a.hi[1], a.hi[2]), o, inv, tmin, best, ta);
          bool hb = slab_test(vec3(b.lo[0], b.lo[1], b.lo[2]), vec3(b.hi[0], b.hi[1], b.hi[2]), o, inv, tmin, best, tb);
          if (ha && hb) {
            bool aFirst = ta <= tb;
            stack[sp] = aFirst ? nd.right : nd.left; entry[sp++] = aFirst ? tb : ta;
            stack[sp] = aFirst ? nd.left : nd.right; entry[sp++] = aFirst ? ta : tb;
          } else if (ha) {
            stack[sp] = nd.left; entry[sp++] = ta;
          } else if (hb) {
            stack[sp] = nd.right; entry[sp++] = tb;
          }
        }
      }
    }
    h.tri = bid;
    h.t = bid == NO_HIT_ID ? 0.f : best;
    h.b1 = bb1;
    h.b2 = bb2;
    return bid != NO_HIT_ID;
  }
};

struct Film {
  uint32_t w = 0, h = 0;
  std::vector<float> accum;   // 2 layers x npix x 3
  std::vector<float> merged;  // npix x 3
  std::vector<uint32_t> counts;
  uint32_t samples = 0;
};

// mergeNewSamples.comp:173-265 for one pass, clamp against the film before this pass
void merge_pass(Film& F, const std::vector<float>& sampleXyz, const ypt_frame_desc& fd) {
  size_t npix = (size_t)F.w * F.h;
  std::vector<float> prevMerged = F.merged;
  std::vector<uint32_t> prevCounts = F.counts;
  for (uint32_t y = 0; y < F.h; ++y)
    for (uint32_t x = 0; x < F.w; ++x) {
      size_t p = (size_t)y * F.w + x;
      float c[3] = {sampleXyz[3 * p], sampleXyz[3 * p + 1], sampleXyz[3 * p + 2]};
      float alpha = 1.f;
      if (fd.kill_nan_samples && (c[0] != c[0] || c[1] != c[1] || c[2] != c[2])) { c[0] = c[1] = c[2] = 0.f; alpha = 0.f; }
      uint32_t newS = (uint32_t)alpha;
      uint32_t oldS = prevCounts[p];
      if (newS > 0) {
        float scale = fd.firefly_clamp_scale;
        if (!(scale <= 0.f || oldS == 0) && !(c[0] <= 0.f && c[1] <= 0.f && c[2] <= 0.f)) {
          float ref[3] = {0.f, 0.f, 0.f};
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              int lx = (int)x + dx, ly = (int)y + dy;
              if (lx < 0 || ly < 0 || lx >= (int)F.w || ly >= (int)F.h) continue;
              size_t q = (size_t)ly * F.w + lx;
              if (prevCounts[q] == 0) continue;
              for (int k = 0; k < 3; ++k) {
                float m = prevMerged[3 * q + k] > 0.f ? prevMerged[3 * q + k] : 0.f;
                ref[k] = ref[k] > m ? ref[k] : m;
              }
            }
          float lvl = ref[0] > ref[1] ? ref[0] : ref[1];
          lvl = lvl > ref[2] ? lvl : ref[2];
          if (lvl > 0.f) {
            for (int k = 0; k < 3; ++k) ref[k] = ref[k] > lvl * 0.05f ? ref[k] : lvl * 0.05f;
            float relax = ypow((float)oldS > 1.f ? (float)oldS : 1.f, fd.firefly_clamp_relax_power);
            float s = 1.f;
            for (int k = 0; k < 3; ++k) {
              float allowed = ref[k] * scale * relax;
              if (c[k] > allowed) s = s < allowed / c[k] ? s : allowed / c[k];
            }
            for (int k = 0; k < 3; ++k) c[k] = c[k] * s;
          }
        }
        F.counts[p] = newS + oldS;
        uint32_t layer = F.samples % 2;
        uint32_t old = oldS / 2;
        double total = (double)(newS + old);
        double db = old / total, sb = newS / total;
        float* a = &F.accum[(layer * npix + p) * 3];
        for (int k = 0; k < 3; ++k) a[k] = (float)(a[k] * db + sb * c[k]);
      }
      float sum[3] = {F.accum[p * 3], F.accum[p * 3 + 1], F.accum[p * 3 + 2]};
      uint32_t targets = F.samples + 1 < 2 ? F.samples + 1 : 2;
      for (uint32_t k = 1; k < targets; ++k)
        for (int j = 0; j < 3; ++j) sum[j] = sum[j] + F.accum[(k * npix + p) * 3 + j];
      for (int j = 0; j < 3; ++j) F.merged[3 * p + j] = sum[j] / (float)targets;
    }
  F.samples++;
}

struct Ctx {
  PreparedScene scene;
  Geometry geo;
  FrameConstants fc;
  SceneView sv;
  bool nee = true;
};

void make_view(Ctx& c) {
  Tables& T = tables();
  c.sv.pos = c.scene.pos.data();
  c.sv.nrm = c.scene.nrm.data();
  c.sv.tan = c.scene.tan.empty() ? nullptr : c.scene.tan.data();
  c.sv.idx = c.scene.idx.data();
  c.sv.triInst = c.scene.triInst.data();
  c.sv.inst = c.scene.inst.data();
  c.sv.mats = c.scene.mats.data();
  c.sv.lights = c.scene.lights.empty() ? nullptr : c.scene.lights.data();
  c.sv.rgb2spec = T.rgb2specFull.empty() ? nullptr : T.rgb2specFull.data();
  c.sv.uv = c.scene.uv.empty() ? nullptr : c.scene.uv.data();
  c.sv.textures = c.scene.textures.empty() ? nullptr : c.scene.textures.data();
  c.sv.texels = c.scene.texels.empty() ? nullptr : c.scene.texels.data();
  c.sv.srgbLut = c.scene.srgbLut.empty() ? nullptr : c.scene.srgbLut.data();
  c.sv.triAlpha = c.scene.alphaTris.empty() ? nullptr : c.scene.triAlpha.data();
  c.sv.alpha = AlphaView{c.scene.alphaTris.empty() ? nullptr : c.scene.alphaTris.data(), c.scene.textures.data(),
                         c.scene.texels.data()};
  c.sv.env = EnvView{c.scene.envTexels.empty() ? nullptr : c.scene.envTexels.data(),
                     c.scene.envTexels.empty() ? ENV_NONE : c.scene.envType, c.scene.envWidth, c.scene.envHeight};
  c.sv.ntris = (uint32_t)c.scene.triInst.size();
  c.sv.ninst = (uint32_t)c.scene.inst.size();
  c.sv.spec.cie_xyz4 = YPT_CIE_XYZ;
  c.sv.spec.d65 = YPT_CIE_D65;
  c.sv.luts.base = T.msl.data();
}

// subsurface probes: the closest hit among one instance's triangles
struct CpuProbe {
  static constexpr bool enabled = true;
  const Geometry* geo;
  uint64_t* work;
  bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h) {
    if (work) work[2]++;
    return geo->closest(o, d, tmin, tmax, h, work, first, count);
  }
};

// rayGenPrimaryRays with rayHitDefault / rayMissEnvironment / inline shadow query
vec3 trace_pixel(const Ctx& c, const uint32_t* point, uint32_t px, uint32_t py, float& lambdaOut, uint64_t* work,
                 FirstHitFeatures* fh) {
  PathState ps;
  start_path(c.fc, point, px, py, ps);
  for (uint32_t i = 0; i < c.fc.maxRayDepth; ++i) {
    if (ps.rayState != RAY_ALIVE) break;
    Hit h;
    c.geo.closest(ps.origin, ps.dir, RAY_MIN_T, __builtin_inff(), h, work);
    if (work) work[2]++;
    ShadowRay sr, sr2;
    sr.valid = false;
    sr2.valid = false;
    CpuProbe probe{&c.geo, work};
    if (h.tri == NO_HIT_ID) shade_miss(c.sv, c.fc, ps);
    else shade_hit(c.sv, c.fc, point, c.nee, ps, h, sr, i == 0 ? fh : nullptr, &probe, &sr2);
    for (const ShadowRay* s : {&sr, &sr2}) {  // the vertex's light connection, then the subsurface exit's
      if (!s->valid) continue;
      Hit sh;
      c.geo.closest(s->origin, s->dir, RAY_MIN_T, s->tmax, sh, work);
      if (work) work[3]++;
      if (sh.tri == s->targetTri) ps.totalLight = ps.totalLight + s->contrib;
    }
    after_segment(point, ps);
  }
  lambdaOut = ps.lambda;
  return path_xyz(c.sv, ps);
}

// ---------------------------------------------------------------- bidirectional integrator
struct CpuTracer {
  const Geometry* geo;
  uint64_t* work;
  bool closest(vec3 o, vec3 d, float tmin, float tmax, Hit& h) {
    if (work) work[2]++;
    return geo->closest(o, d, tmin, tmax, h, work);
  }
  bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h) {
    if (work) work[2]++;
    return geo->closest(o, d, tmin, tmax, h, work, first, count);
  }
  bool hits_target(vec3 o, vec3 d, float tmin, float tmax, uint32_t tri) {
    if (work) work[3]++;
    Hit h;
    geo->closest(o, d, tmin, tmax, h, work);
    return h.tri == tri;
  }
  bool unoccluded(vec3 o, vec3 d, float tmin, float tmax) {
    if (work) work[3]++;
    Hit h;
    return !geo->closest(o, d, tmin, tmax, h, work);
  }
};

template <class F>
void parallel_for(uint32_t n, unsigned nt, F&& f) {
  std::atomic<uint32_t> next{0};
  std::vector<std::thread> th;
  for (unsigned ti = 0; ti < nt; ++ti)
    th.emplace_back([&, ti]() {
      while (true) {
        uint32_t i = next.fetch_add(1);
        if (i >= n) break;
        f(i, ti);
      }
    });
  for (auto& x : th) x.join();
}

// One BDPT pass in the order of the reference's dispatches: light subpaths, camera subpaths per 8x8
// workgroup of permuted texels, light tracing onto the lens, then the per-texel sum of those splats
// in light-subpath order. Only every `stride`-th workgroup is traced when stride > 1.
size_t bdpt_pass(const Ctx& c, const uint32_t* point, const std::vector<uint32_t>& perm, unsigned nt, uint32_t stride,
                 std::vector<float>& xyz, std::vector<float>& lam, std::vector<FirstHitFeatures>* fh,
                 std::vector<std::array<uint64_t, 4>>* work) {
  const FrameConstants& fc = c.fc;
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1], npix = W * H;
  uint32_t npaths = fc.lightPathsPerDim[0] * fc.lightPathsPerDim[1];
  std::vector<LightNodePacked> nodes((size_t)npaths * BDPT_MAX_LIGHT_VERTICES);
  std::vector<uint32_t> counts(npaths, 0);
  std::vector<float> lambdas(npaths, 0.f);
  auto wk = [&](unsigned ti) -> uint64_t* { return work ? (*work)[ti].data() : nullptr; };
  parallel_for(npaths, nt, [&](uint32_t p, unsigned ti) {
    CpuTracer tr{&c.geo, wk(ti)};
    counts[p] = trace_light_path(c.sv, fc, point, p, tr, &nodes[(size_t)p * BDPT_MAX_LIGHT_VERTICES], lambdas[p]);
  });
  std::vector<uint32_t> headers;
  for (uint32_t p = 0; p < npaths; ++p)
    if (counts[p] > 0) headers.push_back(p);
  uint32_t hc = (uint32_t)headers.size();
  uint32_t gcx = (W + BDPT_TILE - 1) / BDPT_TILE, gcy = (H + BDPT_TILE - 1) / BDPT_TILE;
  std::atomic<size_t> traced{0};
  parallel_for(gcx * gcy, nt, [&](uint32_t g, unsigned ti) {
    if (stride > 1 && g % stride != 0) return;
    uint32_t gx = g % gcx, gy = g / gcx;
    float lambda;
    const LightNodePacked* nd = nullptr;
    uint32_t count = 0;
    if (hc > 0) {
      uint32_t p = headers[pick_light_path(point, g, hc)];
      lambda = lambdas[p];
      nd = &nodes[(size_t)p * BDPT_MAX_LIGHT_VERTICES];
      count = counts[p];
    } else {
      uint32_t q = perm[(gy * BDPT_TILE) * W + gx * BDPT_TILE];
      lambda = fallback_wavelength(fc, point, q & 0xFFFFu, q >> 16);
    }
    CpuTracer tr{&c.geo, wk(ti)};
    size_t n = 0;
    for (uint32_t t = 0; t < BDPT_TILE * BDPT_TILE; ++t) {
      uint32_t dx = gx * BDPT_TILE + t % BDPT_TILE, dy = gy * BDPT_TILE + t / BDPT_TILE;
      if (dx >= W || dy >= H) continue;
      uint32_t q = perm[dy * W + dx];
      uint32_t px = q & 0xFFFFu, py = q >> 16;
      size_t pix = (size_t)py * W + px;
      FirstHitFeatures* f = fh ? &(*fh)[pix] : nullptr;
      NodeSurfacesOnTheFly surfs;
      Spectrum L = trace_camera_path(c.sv, fc, point, px, py, lambda, nd, count, surfs, hc == 0 && t == 0, tr, f);
      vec3 v = spectrum_to_xyz(L, spectral_lambdas(lambda), c.sv.spec);
      xyz[3 * pix] = v.x; xyz[3 * pix + 1] = v.y; xyz[3 * pix + 2] = v.z;
      lam[pix] = lambda;
      ++n;
    }
    traced += n;
  });
  uint32_t per = fc.lightPathVertices - 1;
  std::vector<SplatRecord> recs((size_t)npaths * per);
  for (auto& r : recs) r.texel = npix;
  std::vector<uint32_t> recCount(npaths, 0);
  parallel_for(hc, nt, [&](uint32_t h, unsigned ti) {
    uint32_t p = headers[h];
    CpuTracer tr{&c.geo, wk(ti)};
    NodeSurfacesOnTheFly surfs;
    splat_light_path(c.sv, fc, point, h, hc, lambdas[p], &nodes[(size_t)p * BDPT_MAX_LIGHT_VERTICES], counts[p], surfs, tr,
                     &recs[(size_t)h * per]);
    recCount[h] = counts[p] > 0 ? counts[p] - 1 : 0;
  });
  std::vector<float> acc((size_t)npix * 3, 0.f);
  std::vector<uint8_t> has(npix, 0);
  for (uint32_t h = 0; h < hc; ++h)
    for (uint32_t j = 0; j < recCount[h]; ++j) {
      const SplatRecord& r = recs[(size_t)h * per + j];
      if (r.texel >= npix) continue;
      acc[3 * r.texel] += r.xyz.x; acc[3 * r.texel + 1] += r.xyz.y; acc[3 * r.texel + 2] += r.xyz.z;
      has[r.texel] = 1;
    }
  for (uint32_t p = 0; p < npix; ++p)
    if (has[p] && (acc[3 * p] != 0.f || acc[3 * p + 1] != 0.f || acc[3 * p + 2] != 0.f))
      for (int k = 0; k < 3; ++k) xyz[3 * p + k] = xyz[3 * p + k] + acc[3 * p + k];
  return traced;
}

int setup(Ctx& c, const ypt_scene_desc* sd, const ypt_frame_desc* fd, int integrator, int useBvh) {
  if (integrator != YPT_INTEGRATOR_NEE && integrator != YPT_INTEGRATOR_BPT && integrator != YPT_INTEGRATOR_BDPT) {
    g_err = "oracle: integrator not supported";
    return YPT_E_INVALID;
  }
  prepare_scene(*sd, tables().r2s, c.scene);
  prepare_frame(*fd, c.scene, (uint32_t)integrator, c.fc);
  bool needTable = (fd->env_intensity_scale != 0.f && !(fd->flags & YPT_FLAG_WHITE_FURNACE)) || c.fc.punctualLightCount > 0 ||
                   !c.scene.textures.empty();
  if (needTable && tables().rgb2specFull.empty())
    tables().r2s.full_table(tables().rgb2specFull, std::max(1u, std::thread::hardware_concurrency()));
  c.nee = integrator == YPT_INTEGRATOR_NEE;
  c.geo.build(c.scene, useBvh != 0);
  make_view(c);
  return YPT_OK;
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_get_ms_luts(float* out) {
  std::memcpy(out, tables().msl.data(), MSL_TOTAL * sizeof(float));
  return MSL_TOTAL;
}

int oracle_sobol_point(uint32_t index, uint32_t* out256) {
  sobol_point(index, out256);
  return YPT_OK;
}

int oracle_sample_dims(uint32_t sample_index, const uint32_t* seeds, const uint32_t* dims, uint32_t n, float* out) {
  uint32_t pt[256];
  sobol_point(sample_index, pt);
  for (uint32_t i = 0; i < n; ++i) out[i] = sample_dim(pt, dims[i], seeds[i]);
  return YPT_OK;
}

// rgb -> (c0, c1, c2) through the same grid lookup the library uses
int oracle_rgb_coeffs(const float* rgb, float* out) {
  tables().r2s.coeffs(rgb, out);
  return YPT_OK;
}

int oracle_trace_closest(const ypt_scene_desc* sd, const float* rays, uint32_t n, int useBvh, float* t, uint32_t* tri, float* b12) {
  try {
    Ctx c;
    prepare_scene(*sd, tables().r2s, c.scene);
    c.geo.build(c.scene, useBvh != 0);
    for (uint32_t i = 0; i < n; ++i) {
      const float* r = rays + (size_t)i * 8;
      Hit h;
      c.geo.closest(vec3(r[0], r[1], r[2]), vec3(r[4], r[5], r[6]), r[3], r[7], h, nullptr);
      t[i] = h.t; tri[i] = h.tri;
      if (b12) { b12[2 * i] = h.tri == NO_HIT_ID ? 0.f : h.b1; b12[2 * i + 1] = h.tri == NO_HIT_ID ? 0.f : h.b2; }
    }
    return YPT_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return YPT_E_INVALID;
  }
}

// Renders passes [sample_begin, sample_begin + sample_count) and merges them into a film.
// Optional outputs: film (W*H*4: X, Y, Z, 1), counts (W*H), last pass samples (W*H*4: X, Y, Z,
// hero wavelength), first-hit features of the last pass (W*H*12). pixel_stride > 1 renders only every
// pixel_stride-th pixel of every pass (a bounded CPU sample; the film is then partial).
// stats[0] = wall seconds, [1] = path samples, [2] = extension rays, [3] = shadow rays,
// [4] = BVH nodes visited, [5] = triangles tested.
int oracle_render(const ypt_scene_desc* sd, const ypt_frame_desc* fd, int integrator, uint32_t sample_begin,
                  uint32_t sample_count, int threads, int useBvh, uint32_t pixel_stride, float* film, uint32_t* counts,
                  float* last, float* features, double* stats) {
  try {
    Ctx c;
    int rc = setup(c, sd, fd, integrator, useBvh);
    if (rc) return rc;
    uint32_t W = fd->width, H = fd->height;
    size_t npix = (size_t)W * H;
    uint32_t stride = pixel_stride > 1 ? pixel_stride : 1;
    size_t traced = (npix + stride - 1) / stride;
    Film F;
    F.w = W; F.h = H;
    F.accum.assign(2 * npix * 3, 0.f);
    F.merged.assign(npix * 3, 0.f);
    F.counts.assign(npix, 0);
    std::vector<float> xyz(npix * 3, 0.f), lam(npix, 0.f);
    std::vector<FirstHitFeatures> fh(features ? npix : 0);
    unsigned nt = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
    const uint32_t TILE = 16;
    uint32_t tilesX = (W + TILE - 1) / TILE, tilesY = (H + TILE - 1) / TILE;
    std::vector<std::array<uint64_t, 4>> work(nt);
    for (auto& w : work) w = {0, 0, 0, 0};
    std::vector<uint32_t> perm;
    size_t bdptTraced = 0;
    if (integrator == YPT_INTEGRATOR_BDPT) {
      perm.resize(npix);
      texel_permutation(W, H, 0x59415054ull, perm.data());
    }
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t k = 0; k < sample_count; ++k) {
      uint32_t point[256];
      sobol_point(sample_begin + k, point);
      if (features)
        for (auto& f : fh) std::memset(&f, 0, sizeof(f));
      if (integrator == YPT_INTEGRATOR_BDPT) {
        bdptTraced += bdpt_pass(c, point, perm, nt, stride, xyz, lam, features ? &fh : nullptr, stats ? &work : nullptr);
        merge_pass(F, xyz, *fd);
        continue;
      }
      std::atomic<uint32_t> next{0};
      std::vector<std::thread> th;
      for (unsigned ti = 0; ti < nt; ++ti)
        th.emplace_back([&, ti]() {
          uint64_t* wk = stats ? work[ti].data() : nullptr;
          while (true) {
            uint32_t tile = next.fetch_add(1);
            if (tile >= tilesX * tilesY) break;
            uint32_t tx = tile % tilesX, ty = tile / tilesX;
            for (uint32_t y = ty * TILE; y < std::min(H, ty * TILE + TILE); ++y)
              for (uint32_t x = tx * TILE; x < std::min(W, tx * TILE + TILE); ++x) {
                size_t p = (size_t)y * W + x;
                if (stride > 1 && (p % stride) != 0) continue;
                float l;
                vec3 v = trace_pixel(c, point, x, y, l, wk, features ? &fh[p] : nullptr);
                xyz[3 * p] = v.x; xyz[3 * p + 1] = v.y; xyz[3 * p + 2] = v.z;
                lam[p] = l;
              }
          }
        });
      for (auto& x : th) x.join();
      merge_pass(F, xyz, *fd);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (film)
      for (size_t p = 0; p < npix; ++p) {
        film[4 * p] = F.merged[3 * p]; film[4 * p + 1] = F.merged[3 * p + 1]; film[4 * p + 2] = F.merged[3 * p + 2]; film[4 * p + 3] = 1.f;
      }
    if (counts) std::memcpy(counts, F.counts.data(), npix * sizeof(uint32_t));
    if (last)
      for (size_t p = 0; p < npix; ++p) {
        last[4 * p] = xyz[3 * p]; last[4 * p + 1] = xyz[3 * p + 1]; last[4 * p + 2] = xyz[3 * p + 2]; last[4 * p + 3] = lam[p];
      }
    if (features)
      for (size_t p = 0; p < npix; ++p) {
        float* f = features + 12 * p;
        const FirstHitFeatures& a = fh[p];
        if (!a.hit) { std::memset(f, 0, 12 * sizeof(float)); continue; }
        f[0] = a.normal.x; f[1] = a.normal.y; f[2] = a.normal.z;
        f[3] = a.albedo.x; f[4] = a.albedo.y; f[5] = a.albedo.z;
        f[6] = a.roughness; f[7] = a.ior; f[8] = a.metalness; f[9] = a.transparency; f[10] = a.depth; f[11] = 1.f;
      }
    if (stats) {
      uint64_t tot[4] = {0, 0, 0, 0};
      for (auto& w : work) for (int k = 0; k < 4; ++k) tot[k] += w[k];
      stats[0] = std::chrono::duration<double>(t1 - t0).count();
      stats[1] = integrator == YPT_INTEGRATOR_BDPT ? (double)bdptTraced : (double)traced * sample_count;
      stats[2] = (double)tot[2];
      stats[3] = (double)tot[3];
      stats[4] = (double)tot[0];
      stats[5] = (double)tot[1];
      stats[6] = (double)nt;
    }
    return YPT_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return YPT_E_INVALID;
  }
}

}  // extern "C"

// ------------------------------------------------------------------ single-function evaluation for tests
// fn selects the function; in: n rows of 8 floats, out: n rows of 4 floats.
//  0 d_ggx(wm.xyz, ax, ay)            1 pdf_wm_ggx(wo.xyz, wm.xyz, ax, ay)
//  2 sample_wm_ggx(wo.xyz, ax, ay, u1, u2) -> wm   3 fresnel_dielectric(eta, cos)
//  4 fresnel_conductor(n, k, cos)     5 ysin  6 ycos  7 yexp  8 ylog  9 ypow(x, y)
// 10 sample_cosine_hemisphere(u1, u2) 11 sigmoid(x) 12 sigmoid_inv(x) 13 hash_pair(a, b) (uint bits)
// 14 owen_scramble_fast(v, seed) (uint bits)  15 uint_to_float01(v) (uint bits in)
// 16 ss_dir(cos, r)  17 ss_avg(r)  18 ssms_dir(eta, cos, r)  19 sheen_dir(cos, r)  20 tr_dir(eta, cos, r)
// 21 emission_profile(cos, n)  22 f32_to_f16(x) (uint bits out)  23 f16_to_f32(h bits)
// 24 accumulate_ri(sum, nom, denom, connectable)  25 pdf_ratio(nom, denom)
// 26 burley_profile_area(r, d)  27 burley_sample_radius(u, d)  28 burley_d(radius, albedo)
// 29 pdf_subsurface_exit(albedo (grey), radius (grey), to.xyz on the plane y = 0 from the origin, normal +y)
// 30 sellmeier_ior(lambda nm, B1, B2, B3, C1, C2, C3)  31 planck_relative(lambda nm, kelvin)
// 32 avg_fresnel(eta)  33 glambda_sheen(r, cos)  34 d_sheen(cos of the microfacet normal, r)
// 35 refractive_index_for_wavelength(A, B, lambda nm) (Cauchy)
// 36 apply_thin_film(A1, A2, nFilm, etaRTop, phaseShift, cosIncident, thickness nm, hero lambda nm) -> 4 lanes
// 37 ssms_avg(eta, r)  38 average_albedo_translucent(eta, r)
extern "C" int oracle_eval(int fn, const float* in, uint32_t n, float* out) {
  MSLuts L;
  L.base = tables().msl.data();
  for (uint32_t i = 0; i < n; ++i) {
    const float* a = in + 8 * (size_t)i;
    float* o = out + 4 * (size_t)i;
    o[0] = o[1] = o[2] = o[3] = 0.f;
    switch (fn) {
      case 0: o[0] = d_ggx(vec3(a[0], a[1], a[2]), a[3], a[4]); break;
      case 1: o[0] = pdf_wm_ggx(vec3(a[0], a[1], a[2]), vec3(a[3], a[4], a[5]), a[6], a[7]); break;
      case 2: { vec3 w = sample_wm_ggx(vec3(a[0], a[1], a[2]), a[3], a[4], a[5], a[6]); o[0] = w.x; o[1] = w.y; o[2] = w.z; break; }
      case 3: o[0] = fresnel_dielectric(a[0], a[1]); break;
      case 4: o[0] = fresnel_conductor(a[0], a[1], a[2]); break;
      case 5: o[0] = ysin(a[0]); break;
      case 6: o[0] = ycos(a[0]); break;
      case 7: o[0] = yexp(a[0]); break;
      case 8: o[0] = ylog(a[0]); break;
      case 9: o[0] = ypow(a[0], a[1]); break;
      case 10: { vec3 w = sample_cosine_hemisphere(a[0], a[1]); o[0] = w.x; o[1] = w.y; o[2] = w.z; break; }
      case 11: o[0] = sigmoid(a[0]); break;
      case 12: o[0] = sigmoid_inv(a[0]); break;
      case 13: { uint32_t r = hash_pair(as_uint(a[0]), as_uint(a[1])); o[0] = as_float(r); break; }
      case 14: { uint32_t r = owen_scramble_fast(as_uint(a[0]), as_uint(a[1])); o[0] = as_float(r); break; }
      case 15: o[0] = uint_to_float01(as_uint(a[0])); break;
      case 16: o[0] = L.ss_dir(a[0], a[1]); break;
      case 17: o[0] = L.ss_avg(a[0]); break;
      case 18: o[0] = L.ssms_dir(a[0], a[1], a[2]); break;
      case 19: o[0] = L.sheen_dir(a[0], a[1]); break;
      case 20: o[0] = directional_albedo_translucent(L, a[0], a[1], a[2]); break;
      case 21: o[0] = emission_profile(a[0], a[1]); break;
      case 22: o[0] = as_float(f32_to_f16(a[0])); break;
      case 23: o[0] = f16_to_f32(as_uint(a[0])); break;
      case 24: o[0] = accumulate_ri(a[0], a[1], a[2], a[3] != 0.f); break;
      case 25: o[0] = pdf_ratio(a[0], a[1]); break;
      case 26: o[0] = burley_profile_area(a[0], a[1]); break;
      case 27: o[0] = burley_sample_radius(a[0], a[1]); break;
      case 28: o[0] = burley_d(a[0], a[1]); break;
      case 29: o[0] = pdf_subsurface_exit(Spectrum(a[0]), Spectrum(a[1]), vec3(0.f), vec3(0.f, 1.f, 0.f), vec3(a[2], a[3], a[4]),
                                          vec3(a[5], a[6], a[7])); break;
      case 30: o[0] = sellmeier_ior(a + 1, a + 4, a[0]); break;
      case 31: o[0] = planck_relative(a[0], a[1]); break;
      case 32: o[0] = avg_fresnel(a[0]); break;
      case 33: o[0] = glambda_sheen(a[0], a[1]); break;
      case 34: { float c = a[0]; o[0] = d_sheen(vec3(std::sqrt(std::max(0.f, 1.f - c * c)), c, 0.f), a[1]); break; }
      case 35: o[0] = refractive_index_for_wavelength(a[0], a[1], a[2]); break;
      case 36: {
        Lambdas lam = spectral_lambdas(a[7]);
        Spectrum r = apply_thin_film(lam, Spectrum(a[0]), Spectrum(a[1]), a[2], a[3], a[4], a[5], a[6]);
        for (int k = 0; k < 4; ++k) o[k] = r.s[k];
        break;
      }
      case 37: o[0] = L.ssms_avg(a[0], a[1]); break;
      case 38: o[0] = average_albedo_translucent(L, a[0], a[1]); break;
      case 39: o[0] = yatan2(a[0], a[1]); break;
      case 40: o[0] = yacos(a[0]); break;
      case 41: o[0] = yasin(a[0]); break;
      default: g_err = "unknown function"; return YPT_E_INVALID;
    }
  }
  return YPT_OK;
}

// the colour that a coefficient triplet reproduces under D65 (SpectralUtility.cpp:455-476), linear sRGB
extern "C" int oracle_rgb_reproduce(const float* coeffs, float* rgb_out) {
  double c[3] = {coeffs[0], coeffs[1], coeffs[2]}, r[3];
  tables().r2s.reproduce(c, r);
  for (int k = 0; k < 3; ++k) rgb_out[k] = (float)r[k];
  return YPT_OK;
}

// evaluate_surface for one untextured material in a frame whose geometry and shading normals are +y
// and whose tangent is +x, seen from outside (current and previous IOR 1): the layered BSDF's
// getSkyBoxColor of the shared core over n directions (3 floats each): RGB per direction, for an
// environment as the scene description takes it (NULL: the texture-less |direction| colour)
extern "C" int oracle_env_color(const ypt_environment* env, const float* dirs, uint32_t n, float* rgb) {
  try {
    ypt_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.environment = env;
    PreparedScene ps;
    prepare_environment(d, ps);
    SceneView sv;
    std::memset(&sv, 0, sizeof(sv));
    sv.env = EnvView{ps.envTexels.empty() ? nullptr : ps.envTexels.data(), ps.envType, ps.envWidth, ps.envHeight};
    for (uint32_t i = 0; i < n; ++i) {
      vec3 c = env_color(sv, vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
      rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z;
    }
    return YPT_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return YPT_E_INVALID;
  }
}

// weight per hero lane (projected by |wi.y| unless flags ask otherwise) and its summed pdf.
// out: weight[4], pdf, transmission flags
extern "C" int oracle_eval_surface(const ypt_material* m, const float* wo, const float* wi, float hero, uint32_t flags,
                                   int reciprocal, float* out6) {
  try {
    Tables& T = tables();
    MaterialGPU mg = convert_material(T.r2s, *m, 0);
    SpectralTables st;
    st.cie_xyz4 = YPT_CIE_XYZ;
    st.d65 = YPT_CIE_D65;
    Lambdas lam = spectral_lambdas(hero);
    vec3 n(0.f, 1.f, 0.f), t(1.f, 0.f, 0.f), o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    SurfaceDef sd = make_surface(mg, n, n, t, o, true, false, lam, st);
    SurfaceContext c;
    make_surface_context(c, sd, 1.f, 1.f, o, true, lam, reciprocal != 0);
    MSLuts L;
    L.base = T.msl.data();
    SurfaceEval e = evaluate_surface(c, L, o, i, flags);
    for (int k = 0; k < 4; ++k) out6[k] = e.weight.s[k];
    out6[4] = e.pdf;
    out6[5] = (float)e.transmission;
    return YPT_OK;
  } catch (const std::exception& e) {
    g_err = e.what();
    return YPT_E_INVALID;
  }
}
