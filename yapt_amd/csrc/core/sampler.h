// Hash-scrambled Sobol sampler.
//
// Restates shaders/common/random.hlsl:35-92 (murmur finaliser, pair hash, fast base-2 Owen
// scramble, uint->[0,1) conversion) and shaders/common/randomSampler.hlsl:23-102 (per-dimension
// scrambling and the RandomSampler dimension bookkeeping), and the host generator
// Source/Common/inc/Math/RandUtility.inl:33-45 (sobolUint) that fills the per-frame point.
//
// Every pixel of a frame reads the same Sobol point (index = frame sample index, 256 dimensions,
// RaytraceCommonResources.cpp:623-627); pixels are decorrelated by scrambling with
// hash(pixelIndex, dimension). Dimensions at or above YPT_SOBOL_POINT_DIMS read 0: the reference
// reads them past the end of a 256-entry raw buffer, which D3D12 defines to return 0.
#pragma once
#include "common.h"

namespace ypt {

static constexpr uint32_t YPT_SOBOL_POINT_DIMS = 256;  // NUMBER_OF_RANDOM_SAMPLE_DIMENSIONS

YPT_HD uint32_t reverse_bits32(uint32_t n) {
  n = (n << 16) | (n >> 16);
  n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
  n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
  n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
  n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
  return n;
}

// Sobol point `index` in `dimension` from 32-row generator matrices (RandUtility.inl:33-45).
YPT_HD uint32_t sobol_uint(const uint32_t* matrices, uint32_t index, uint32_t dimension) {
  uint32_t v = 0;
  for (uint32_t i = dimension * 32u; index != 0; index >>= 1, ++i)
    if (index & 1u) v ^= matrices[i];
  return v;
}

YPT_HD uint32_t hash_murmur_final(uint32_t x) {  // random.hlsl:35-43
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

// random.hlsl:50-54. Integer division, 32-bit wrap-around as in HLSL uint arithmetic.
YPT_HD uint32_t hash_pair(uint32_t a, uint32_t b) {
  uint32_t key = ((a + b) * (a + b + 1u)) / (2u + b);
  return hash_murmur_final(key);
}

YPT_HD uint32_t owen_scramble_fast(uint32_t v, uint32_t seed) {  // random.hlsl:75-84
  v = reverse_bits32(v);
  v ^= v * 0x3d20adeau;
  v += seed;
  v *= (seed >> 16) | 1u;
  v ^= v * 0x05526c56u;
  v ^= v * 0x53a22864u;
  return reverse_bits32(v);
}

// [0, 0xFFFFFFFF] -> [0, 1)   (random.hlsl:88-91)
YPT_HD float uint_to_float01(uint32_t r) {
  float f = (float)r * 2.3283064365386963e-10f;  // 0x1p-32
  return fmin_(f, 0.99999994f);                  // 0x1.fffffep-1
}

YPT_HD uint32_t scramble(uint32_t random, uint32_t seed, uint32_t dimension) {  // randomSampler.hlsl:23-27
  return owen_scramble_fast(random, hash_pair(seed, dimension));
}

YPT_HD uint32_t sobol_point_dim(const uint32_t* point, uint32_t d) {
  return d < YPT_SOBOL_POINT_DIMS ? point[d] : 0u;
}

YPT_HD float sample_dim(const uint32_t* point, uint32_t d, uint32_t seed) {
  return uint_to_float01(scramble(sobol_point_dim(point, d), seed, d));
}

// RandomSampler (randomSampler.hlsl:57-102): x = next dimension, y = scramble seed (pixel or path index).
struct RandomSampler {
  uint32_t dim;
  uint32_t seed;
  const uint32_t* point;

  YPT_HD void init(const uint32_t* p, uint32_t d, uint32_t pixelIndex) {
    point = p;
    dim = d + 2u;  // first two dimensions are the texel offset
    seed = pixelIndex;
  }
  YPT_HD void texel_offset(float& u, float& v) const {
    u = sample_dim(point, 0u, seed);
    v = sample_dim(point, 1u, seed);
  }
  YPT_HD float r1() { float v = sample_dim(point, dim, seed); dim += 1u; return v; }
  YPT_HD void r2(float out[2]) {
    out[0] = sample_dim(point, dim, seed);
    out[1] = sample_dim(point, dim + 1u, seed);
    dim += 2u;
  }
  YPT_HD void r4(float out[4]) {
    for (uint32_t k = 0; k < 4u; ++k) out[k] = sample_dim(point, dim + k, seed);
    dim += 4u;
  }
};

}  // namespace ypt
