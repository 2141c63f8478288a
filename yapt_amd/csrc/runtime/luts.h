// Host generation of the microfacet directional-albedo tables (energy compensation).
//
// Restates Source/Renderer/inc/Renderer/Shared/Utility/MultipleScatterLUTUtility.h:8-438 and the
// table loops of Source/Renderer/src/Shared/Utility/MultipleScatterLUT.cpp:66-380, with the CPU
// microfacet twins they call (Source/Common/inc/Math/GGX.h:12-290, Fresnel.h:9-80, MiscBrdf.h:11-91,
// RandUtility.h:47-60). Differences, both deliberate:
//  * the translucent tables integrate over a Halton sequence with identity digit permutations; the
//    reference shuffles digits with an unseeded CRT rand() (RandUtility.cpp:11-35), so its tables are
//    not reproducible and cannot be matched;
//  * all loops are deterministic and the result does not depend on the thread count.
// The output layout is the MSL_* block of core/msluts.h.
#pragma once
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>
#include <algorithm>

#include "../core/msluts.h"

namespace ypt {
namespace lutgen {

struct v3 {
  float x, y, z;
};
static inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dotp(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 crossp(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline v3 norm(v3 a) { float l = std::sqrt(dotp(a, a)); return mk(a.x / l, a.y / l, a.z / l); }
static inline v3 reflectp(v3 i, v3 n) { return sub(i, scl(n, 2.f * dotp(n, i))); }
static inline v3 refractp(v3 i, v3 n, float eta) {  // glm::refract
  float d = dotp(n, i);
  float k = 1.f - eta * eta * (1.f - d * d);
  if (k < 0.f) return mk(0.f, 0.f, 0.f);
  return sub(scl(i, eta), scl(n, eta * d + std::sqrt(k)));
}
static inline float sqrf(float v) { return v * v; }
static inline float satf(float s) { return std::min(std::max(s, 0.f), 1.f); }
static inline float stepf(float y, float x) { return (x >= y) ? 1.f : 0.f; }
static inline float sdiv(float x, float y) { return y == 0.f ? 0.f : x / y; }
static const float LPI = 3.14159265358979323846f;

static inline float smith_lambda(v3 d, float ax, float ay) {
  float s = 1.0f + (d.x * d.x * ax * ax + d.z * d.z * ay * ay) / (d.y * d.y);
  return (std::sqrt(s) - 1) * 0.5f;
}
static inline float g1(v3 wo, v3 wm, float ax, float ay) { return stepf(0, dotp(wo, wm)) / (1.0f + smith_lambda(wo, ax, ay)); }
static inline float g2(v3 wo, v3 wi, v3 wm, float ax, float ay) {
  float li = smith_lambda(wi, ax, ay), lo = smith_lambda(wo, ax, ay);
  return stepf(0, dotp(wo, wm)) / (1 + li + lo);
}
static inline float dggx(v3 wm, float ax, float ay) {
  ax = std::max(ax, 0.000001f);
  ay = std::max(ay, 0.000001f);
  float denom = sqrf(sdiv(wm.x, ax)) + sqrf(sdiv(wm.z, ay)) + sqrf(wm.y);
  denom *= denom;
  denom *= ax * ay * LPI;
  return sdiv(1.f, denom);
}
static inline float dvggx(v3 wo, v3 wm, float ax, float ay) {
  return sdiv(g1(wo, wm, ax, ay) * std::fabs(dotp(wo, wm)) * dggx(wm, ax, ay), std::fabs(wo.y));
}
static inline v3 sample_vn(v3 wo, float ax, float ay, float u1, float u2) {
  v3 Vh = norm(mk(ax * wo.x, ay * wo.z, wo.y));
  float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
  v3 T1 = lensq > 0 ? scl(mk(-Vh.y, Vh.x, 0), 1.f / std::sqrt(lensq)) : mk(1, 0, 0);
  v3 T2 = crossp(Vh, T1);
  float r = std::sqrt(u1);
  float phi = 2.0f * LPI * u2;
  float t1 = r * std::cos(phi);
  float t2 = r * std::sin(phi);
  float s = 0.5f * (1.0f + Vh.z);
  t2 = (1.0f - s) * std::sqrt(1.0f - t1 * t1) + s * t2;
  v3 Nh = add(add(scl(T1, t1), scl(T2, t2)), scl(Vh, std::sqrt(std::max(0.f, 1.f - t1 * t1 - t2 * t2))));
  return norm(mk(ax * Nh.x, std::max(0.0f, Nh.z), ay * Nh.y));
}
static inline v3 sample_wm(v3 wo, float ax, float ay, float u1, float u2) {
  v3 wh = sample_vn(wo, ax, ay, u1, u2);
  if (wo.y * wh.y < 0) wh = neg(wh);
  return wh;
}
static inline float jrefl(v3 wi, v3 wm) { return sdiv(1.f, 4 * satf(dotp(wi, wm))); }
static inline float jrefr(float eta, v3 wo, v3 wm, v3 wi) {
  float dotMO = dotp(wo, wm);
  float dotMI = std::fabs(dotp(wi, wm));
  float denom = sqrf(dotMO + eta * dotMI);
  return sdiv(sqrf(eta) * dotMI, denom);
}
static inline float fresnel_diel(float eta, float cosTheta) {
  float c = satf(cosTheta);
  float temp = eta * eta + c * c - 1.f;
  if (temp < 0.f) return 1.f;
  float g = std::sqrt(temp);
  return 0.5f * sqrf((g - c) / std::max(g + c, 0.000001f)) * (1 + sqrf(((g + c) * c - 1.f) / ((g - c) * c + 1.f)));
}
static inline v3 wm_translucent(v3 wo, v3 wi, float etaR) {
  if (wi.y >= 0.f) return norm(add(wo, wi));
  if (etaR == 1.f) etaR = 1.0001f;
  v3 wm = norm(add(wo, scl(wi, etaR)));
  if (wo.y * wm.y < 0) wm = scl(wm, -1.f);
  return wm;
}
static inline v3 sample_translucent(float etaR, float ax, float ay, v3 wo, v3 s) {
  v3 wm = sample_wm(wo, ax, ay, s.x, s.y);
  if (etaR == 1.f) etaR = 1.0001f;
  float F = fresnel_diel(etaR, dotp(wo, wm));
  if (dotp(wo, wm) < 0.f) return mk(0, 0, 0);
  float invEta = 1.f / etaR;
  v3 wi;
  if (s.z < F) {
    wi = reflectp(neg(wo), wm);
    if (wi.y <= 0.f) wi = mk(0, 0, 0);
  } else {
    wi = refractp(neg(wo), wm, invEta);
    if (wi.y > 0) wi = mk(0, 0, 0);
  }
  return wi;
}
static inline float pdf_translucent(float etaR, v3 wo, v3 wi, float ax, float ay) {
  bool refl = wi.y >= 0.f;
  v3 wm = wm_translucent(wo, wi, etaR);
  if (dotp(wo, wm) <= 0.f) return 0.f;
  float F = fresnel_diel(etaR, dotp(wo, wm));
  float pdf = dvggx(wo, wm, ax, ay);
  if (refl) pdf *= F * jrefl(wo, wm);
  else pdf *= (1.f - F) * jrefr(etaR, wo, wm, wi);
  return pdf;
}
static inline float eval_translucent(float etaR, float ax, float ay, v3 wo, v3 wi) {
  bool refl = wi.y >= 0.f;
  v3 wm = wm_translucent(wo, wi, etaR);
  float VdotH = satf(dotp(wo, wm));
  float F = fresnel_diel(etaR, VdotH);
  float G2 = g2(wo, wi, wm, ax, ay);
  float D = dggx(wm, ax, ay);
  if (refl) return F * G2 * D / (4.f * wo.y * wi.y);
  if (dotp(wo, wm) * dotp(wi, wm) > 0) return 0.f;
  float denom = std::fabs(wo.y * wi.y);
  float w = (1.f - F) * G2 * D * VdotH * jrefr(etaR, wo, wm, wi);
  return sdiv(w, denom);
}

static inline uint32_t rev32(uint32_t n) {
  n = (n << 16) | (n >> 16);
  n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
  n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
  n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
  n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
  return n;
}
static inline void hammersley(uint32_t i, uint32_t n, float& a, float& b) {
  const float invMax = (float)(1.0 / 4294967295.0);
  a = float(rev32(i)) * invMax;
  b = float(i) / float(n);
}
// radical inverse (pbrt halton with identity digit permutation), RandUtility.inl:12-30
static inline float halton(uint64_t base, uint64_t index) {
  uint64_t limit = ~0ull / base - base;
  float invBase = 1.f / (float)base;
  float invBaseM = 1.f;
  uint64_t reversed = 0;
  while (1 - (base - 1) * invBaseM < 1 && reversed < limit) {
    uint64_t next = index / base;
    uint64_t digit = index - next * base;
    reversed = reversed * base + digit;
    invBaseM *= invBase;
    index = next;
  }
  return std::min(invBaseM * reversed, 0x1.fffffep-1f);
}

// CPU sampleLut variants (MultipleScatterLUTUtility.h:36-92): nearest-floor/ceil linear, no half-texel remap
static inline float slut1(float u, const float* lut, uint32_t dim) {
  u = satf(u);
  float v = u * dim;
  v = std::min(v, float(dim - 1));
  float a = lut[uint32_t(std::floor(v))], b = lut[uint32_t(std::ceil(v))];
  float t = v - std::floor(v);
  return (1.f - t) * a + t * b;
}
static inline float slut2(float ux, float uy, const float* lut, uint32_t dx, uint32_t dy) {
  ux = satf(ux); uy = satf(uy);
  float vx = std::min(ux * dx, float(dx - 1)), vy = std::min(uy * dy, float(dy - 1));
  auto at = [&](float x, float y) { return lut[size_t(y * dx + x)]; };
  float a = at(std::floor(vx), std::floor(vy)), b = at(std::ceil(vx), std::floor(vy));
  float c = at(std::floor(vx), std::ceil(vy)), d = at(std::ceil(vx), std::ceil(vy));
  float tx = vx - std::floor(vx), ty = vy - std::floor(vy);
  float ab = (1.f - tx) * a + tx * b, cd = (1.f - tx) * c + tx * d;
  return (1.f - ty) * ab + ty * cd;
}
static inline float slut3(float ux, float uy, float uz, const float* lut, uint32_t d) {
  ux = satf(ux); uy = satf(uy); uz = satf(uz);
  float vz = std::min(uz * d, float(d - 1));
  size_t z = size_t(std::floor(vz));
  float a = slut2(ux, uy, lut + z * d * d, d, d);
  z = std::min(z + 1, size_t(d - 1));
  float b = slut2(ux, uy, lut + z * d * d, d, d);
  float t = vz - std::floor(vz);
  return (1.f - t) * a + t * b;
}
static inline float avg_albedo(const float* dir, uint32_t dim) {  // integrateAverageAlbedo
  float h = 1.f / float(dim - 1);
  float average = 0.f;
  for (uint32_t x = 0; x + 1 < dim; ++x) {
    float a = float(x) * h, b = float(x + 1) * h;
    average += h * (dir[x] * (2.f * a + b) + dir[x + 1] * (a + 2.f * b)) / 6.f;
  }
  return 2.f * average;
}
static inline float valid_eta(float etaR) { return etaR > 1 ? std::max(etaR, 1.01f) : std::min(etaR, 0.99f); }
static inline float avg_fresnel_cpu(float etaR) {
  if (etaR < 1) return 0.997118f + 0.1014f * etaR - 0.965241f * etaR * etaR - 0.130607f * etaR * etaR * etaR;
  return (etaR - 1.f) / (4.08567f + 1.00071f * etaR);
}

static inline float dir_ss_nofresnel(float roughness, float cosTheta) {
  const uint32_t N = 256;
  float accum = 0;
  roughness = std::max(0.01f, roughness);
  float a = roughness * roughness;
  float sinTheta = std::sqrt(1.f - sqrf(cosTheta));
  v3 wo = mk(sinTheta, cosTheta, 0);
  for (uint32_t i = 0; i < N; ++i) {
    float s0, s1;
    hammersley(i, N, s0, s1);
    v3 wm = sample_wm(wo, a, a, s0, s1);
    v3 wi = reflectp(neg(wo), wm);
    if (wi.y <= 0.0f || dotp(wi, wm) <= 0.f || dotp(wo, wm) <= 0.f) continue;
    float G2 = g2(wo, wi, wm, a, a), D = dggx(wm, a, a);
    float pdf = dvggx(wo, wm, a, a) * jrefl(wo, wm);
    float w = G2 * D / std::max(4.f * wo.y * wi.y, 0.000001f);
    if (pdf > 0.f) accum += w * std::fabs(wi.y) / pdf;
  }
  return accum / N;
}
static inline float kulla_cpu(float fms, float dotWo, float dotWi, float lr, const float* ss, const float* ssAvg) {
  dotWo = satf(dotWo); dotWi = satf(dotWi);
  float eo = 1.f - slut2(dotWo, lr, ss, SS_DIM, SS_DIM);
  float ei = 1.f - slut2(dotWi, lr, ss, SS_DIM, SS_DIM);
  float eAvg = slut1(lr, ssAvg, SS_DIM);
  float ems = eo * ei / std::max(0.0001f, (1.f - eAvg) * LPI);
  return fms * ems;
}
static inline float dir_ssms(float roughness, float cosTheta, float etaR, const float* ss, const float* ssAvg) {
  const uint32_t N = 256;
  float accum = 0;
  roughness = std::max(0.01f, roughness);
  float a = roughness * roughness;
  etaR = valid_eta(etaR);
  float sinTheta = std::sqrt(1.f - sqrf(cosTheta));
  v3 wo = mk(sinTheta, cosTheta, 0);
  for (uint32_t i = 0; i < N; ++i) {
    float s0, s1;
    hammersley(i, N, s0, s1);
    v3 wm = sample_wm(wo, a, a, s0, s1);
    v3 wi = reflectp(neg(wo), wm);
    if (wi.y < 0.0f || dotp(wi, wm) < 0.f || dotp(wo, wm) < 0.f) continue;
    float G2 = g2(wo, wi, wm, a, a);
    float F = fresnel_diel(etaR, dotp(wo, wm));
    float D = dggx(wm, a, a);
    float pdf = dvggx(wo, wm, a, a) * jrefl(wo, wm);
    float w = F * G2 * D / std::max(4.f * wo.y * wi.y, 0.000001f);
    w += kulla_cpu(F, cosTheta, wi.y, roughness, ss, ssAvg);
    if (pdf > 0.f) accum += w * std::fabs(wi.y) / pdf;
  }
  return accum / N;
}
static inline float dir_translucent(float etaR, float roughness, float cosTheta) {
  const uint32_t N = 1024;
  float accum = 0;
  etaR = valid_eta(etaR);
  roughness = std::max(0.01f, roughness);
  float a = roughness * roughness;
  float sinTheta = std::sqrt(1.f - sqrf(cosTheta));
  v3 wo = mk(sinTheta, cosTheta, 0);
  for (uint32_t i = 0; i < N; ++i) {
    v3 s = mk(halton(2, i), halton(3, i), halton(5, i));
    v3 wi = sample_translucent(etaR, a, a, wo, s);
    if (dotp(wi, wi) == 0.f) continue;
    float pdf = pdf_translucent(etaR, wo, wi, a, a);
    float w = eval_translucent(etaR, a, a, wo, wi);
    if (pdf > 0.f) accum += w * std::fabs(wi.y) / pdf;
  }
  return accum / N;
}
static inline float glsheen(float x, float a, float b, float c, float d, float e) { return (a / (1 + b * std::pow(x, c))) + d * x + e; }
static inline float lerpc(float x, float y, float t) { return x * (1.f - t) + y * t; }
static inline float glambda_sheen(float r, float cosTheta) {
  float a0 = 25.3245f, b0 = 3.32435f, c0 = 0.16801f, d0 = -1.27393f, e0 = -4.85967f;
  float a1 = 21.5473f, b1 = 3.82987f, c1 = 0.19823f, d1 = -1.97760f, e1 = -4.32054f;
  float t = 1.f - r;
  t *= t;
  float a = lerpc(a1, a0, t), b = lerpc(b1, b0, t), c = lerpc(c1, c0, t), d = lerpc(d1, d0, t), e = lerpc(e1, e0, t);
  float ex = cosTheta < 0.5f ? glsheen(cosTheta, a, b, c, d, e) : 2.f * glsheen(0.5f, a, b, c, d, e) - glsheen(1.f - cosTheta, a, b, c, d, e);
  return std::exp(ex);
}
static inline float dir_sheen(float roughness, float cosTheta) {
  const uint32_t N = 256;
  float accum = 0;
  roughness = std::max(0.01f, roughness);
  float a = std::max(roughness, 0.07f);
  float sinTheta = std::sqrt(1.f - sqrf(cosTheta));
  v3 wo = mk(sinTheta, cosTheta, 0);
  for (uint32_t i = 0; i < N; ++i) {
    float s0, s1;
    hammersley(i, N, s0, s1);
    float z = s0, r = std::sqrt(1.f - z * z), phi = 2.f * LPI * s1;
    v3 wi = mk(r * std::cos(phi), z, r * std::sin(phi));
    v3 wm = norm(add(wi, wo));
    if (wi.y <= 0.0f) continue;
    float G = 1.f / (1.f + glambda_sheen(a, wi.y) + glambda_sheen(a, wo.y));
    float rInv = 1.f / a;
    float D = (2.f + rInv) * std::pow(1.f - wm.y * wm.y, rInv * 0.5f) / (2.f * LPI);
    float pdf = 1.f / (2 * LPI);
    float w = G * D / std::max(4.f * wo.y * wi.y, 0.000001f);
    if (pdf > 0.f) accum += w * std::fabs(wi.y) / pdf;
  }
  return accum / N;
}

template <typename F>
static void parallel_for(int n, F f) {
  unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t]() { for (int i = (int)t; i < n; i += (int)nt) f(i); });
  for (auto& x : th) x.join();
}

}  // namespace lutgen

// Fills the MSL_TOTAL-float block (core/msluts.h layout).
inline void generate_ms_luts(float* out) {
  using namespace lutgen;
  const float MIN_COS = 0.001f;
  float* ss = out + MSL_SS_DIR;
  float* ssAvg = out + MSL_SS_AVG;
  parallel_for(SS_DIM, [&](int y) {
    for (int x = 0; x < SS_DIM; ++x) {
      float c = float(x) / (SS_DIM - 1), r = float(y) / (SS_DIM - 1);
      ss[y * SS_DIM + x] = dir_ss_nofresnel(r, std::max(c, MIN_COS));
    }
  });
  for (int y = 0; y < SS_DIM; ++y) ssAvg[y] = avg_albedo(ss + y * SS_DIM, SS_DIM);

  float* ssms = out + MSL_SSMS_DIR;
  float* ssmsAvg = out + MSL_SSMS_AVG;
  parallel_for(SSMS_DIM, [&](int z) {
    for (int y = 0; y < SSMS_DIM; ++y)
      for (int x = 0; x < SSMS_DIM; ++x) {
        float r = float(y) / (SSMS_DIM - 1), c = float(x) / (SSMS_DIM - 1);
        float eta = float(z) / (SSMS_DIM - 1);
        eta = 1.f + eta * 3.f;
        eta = std::max(eta, 1.01f);
        ssms[(z * SSMS_DIM + y) * SSMS_DIM + x] = dir_ssms(r, std::max(c, MIN_COS), eta, ss, ssAvg);
      }
  });
  for (int z = 0; z < SSMS_DIM; ++z)
    for (int y = 0; y < SSMS_DIM; ++y) ssmsAvg[z * SSMS_DIM + y] = avg_albedo(ssms + (z * SSMS_DIM + y) * SSMS_DIM, SSMS_DIM);

  struct Cfg { int dir, avg; float scale, bias, mn, mx; };
  const Cfg cfgs[2] = {{MSL_TR_DIR_DENSER, MSL_TR_AVG_DENSER, 3.f, 1.f, 1.0001f, 4.0f},
                       {MSL_TR_DIR_LIGHTER, MSL_TR_AVG_LIGHTER, 0.69f, 0.3f, 0.3f, 0.99f}};
  for (const Cfg& cf : cfgs) {
    float* d = out + cf.dir;
    parallel_for(TR_DIM, [&](int z) {
      for (int y = 0; y < TR_DIM; ++y)
        for (int x = 0; x < TR_DIM; ++x) {
          float r = float(y) / (TR_DIM - 1), c = float(x) / (TR_DIM - 1);
          float eta = float(z) / (TR_DIM - 1);
          eta = cf.scale * eta + cf.bias;
          eta = std::min(std::max(eta, cf.mn), cf.mx);
          d[(z * TR_DIM + y) * TR_DIM + x] = dir_translucent(eta, r, std::max(c, MIN_COS));
        }
    });
    for (int z = 0; z < TR_DIM; ++z)
      for (int y = 0; y < TR_DIM; ++y) out[cf.avg + z * TR_DIM + y] = avg_albedo(d + (z * TR_DIM + y) * TR_DIM, TR_DIM);
  }

  float* sh = out + MSL_SHEEN_DIR;
  parallel_for(SHEEN_DIM, [&](int y) {
    for (int x = 0; x < SHEEN_DIM; ++x) {
      float c = float(x) / (SHEEN_DIM - 1), r = float(y) / (SHEEN_DIM - 1);
      sh[y * SHEEN_DIM + x] = dir_sheen(r, std::max(c, MIN_COS));
    }
  });
}

}  // namespace ypt
