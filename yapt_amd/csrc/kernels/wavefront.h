// Wavefront kernels of the unidirectional integrators (BPT, BPT_WITH_NEE) for gfx950.
//
// The reference runs each path start-to-finish in one DXR ray-generation thread
// (rayGenPrimaryRays.hlsl:10-145). Here a pass over one Sobol sample index is split into kernels
// that each do one kind of work for all live paths, with path state in HBM as structure-of-arrays:
//   start   - camera rays for every pixel, ray queue = all pixels
//   extend  - closest-hit traversal of the ray queue (hit records per path slot)
//   shade   - surface/miss evaluation, light sample, next direction, Russian roulette; appends
//             shadow rays and surviving paths to two compacted queues (wave ballot + one atomic
//             per wave)
//   shadow  - occlusion queries of the shadow queue, adds visible light contributions
//   merge   - XYZ of each path, film running mean with the firefly clamp (mergeNewSamples.comp)
// Queue order is whatever the atomics produce; no result depends on it because every path only
// touches its own slot.
#pragma once
#include "../core/shade.h"
#include "traverse.h"
#include "traverse_wave.h"

namespace ypt {

struct PathSoA {
  float4* orgLambda;   // origin xyz, hero wavelength
  float4* dirPdf;      // direction xyz, pdfThisRay
  float4* thr;         // throughput
  float4* light;       // totalLight
  uint4* st0;          // dim, seed, flags, pathLength
  uint4* st1;          // rayState, nvol, furnaceInstance, -
  float4* volIor;      // 4 iors
  uint4* volMat;       // 4 material ids
  float4* volAbs;      // 4 spectra per slot: [slot * 4 + k]
};

struct Queues {
  uint32_t* ray[2];
  uint32_t* counters;  // [0], [1]: ray queue lengths, [2]: shadow queue length, [3..]: stats
  float4* hit;         // per slot: t, b1, b2, tri (bits)
  float4* sOrgTmax;    // shadow queue
  float4* sDirTri;
  float4* sContrib;
  uint32_t* sSlot;
  uint32_t* ovf;       // traversal stack overflow
  uint32_t ovfStride;
  unsigned long long* workCounters;  // [0] nodes, [1] tris (extension), [2] nodes, [3] tris (shadow)
  uint32_t* sssStack;  // subsurface probe traversal stacks of the shading launch, [depth][threads]
  uint32_t waveBelow;  // ray and shadow queues shorter than this are traced wave-per-ray (k_extend_w, k_shadow_w)
  uint32_t stepsPerFetch;  // persistent traversal: node steps between a wave's refill checks (>= 1)
};

struct FilmBuffers {
  float4* accum;       // [2][npix]
  float4* merged[2];   // double buffer: read previous, write current
  uint32_t* counts[2];
  float* variance;
  float4* lastSample;  // X, Y, Z, hero wavelength
  float* features;     // FEATURE_STRIDE floats per path slot (first-hit features), optional
  float4* featAccum;   // [4][npix] running means of the feature moments, with features
};

// featureSampleToMoments (denoiseFeatures.hlsl:29-44): albedo, |albedo|^2 / normal, coverage /
// roughness, metalness, transparency, anisotropy / ior * 0.5, |material|^2, depth, depth^2; zeros for
// a miss or a non-finite record
YPT_HD void feature_moments(const float* f, float4 m[4]) {
  bool finite = true;
  for (int k = 0; k < 13; ++k) finite = finite && !is_nan(f[k]) && fabsf(f[k]) <= 3.402823466e38f;
  if (!(f[11] > 0.f) || !finite) {
    for (int k = 0; k < 4; ++k) m[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  float ax = f[3], ay = f[4], az = f[5];
  float r = f[6], me = f[8], tr = f[9], an = f[12];
  float ior = f[7] * 0.5f;  // FEATURE_IOR_SCALE
  float d = f[10];
  m[0] = make_float4(ax, ay, az, ax * ax + ay * ay + az * az);
  m[1] = make_float4(f[0], f[1], f[2], 1.f);
  m[2] = make_float4(r, me, tr, an);
  m[3] = make_float4(ior, (r * r + me * me + tr * tr + an * an) + ior * ior, d, d * d);
}

struct DevScene {
  SceneView sv;
  const Bvh2Node* nodes2;
  const Bvh8Node* nodes8;
  TriSet tris;
};

__device__ __forceinline__ void load_path(const PathSoA& P, uint32_t s, PathState& ps) {
  float4 a = P.orgLambda[s], b = P.dirPdf[s], t = P.thr[s], l = P.light[s];
  uint4 u0 = P.st0[s], u1 = P.st1[s];
  ps.origin = vec3(a.x, a.y, a.z);
  ps.lambda = a.w;
  ps.dir = vec3(b.x, b.y, b.z);
  ps.pdfThisRay = b.w;
  ps.throughput = Spectrum(t.x, t.y, t.z, t.w);
  ps.totalLight = Spectrum(l.x, l.y, l.z, l.w);
  ps.dim = u0.x; ps.seed = u0.y; ps.flags = u0.z; ps.pathLength = u0.w;
  ps.rayState = u1.x; ps.nvol = u1.y; ps.furnaceInstance = u1.z;
  if (ps.nvol > 0) {
    float4 io = P.volIor[s];
    uint4 vm = P.volMat[s];
    ps.volIor[0] = io.x; ps.volIor[1] = io.y; ps.volIor[2] = io.z; ps.volIor[3] = io.w;
    ps.volMat[0] = vm.x; ps.volMat[1] = vm.y; ps.volMat[2] = vm.z; ps.volMat[3] = vm.w;
    for (int k = 0; k < 4; ++k) {
      float4 ab = P.volAbs[(size_t)s * 4 + k];
      ps.volAbs[k] = Spectrum(ab.x, ab.y, ab.z, ab.w);
    }
  } else {
    for (int k = 0; k < 4; ++k) { ps.volIor[k] = 1.f; ps.volMat[k] = VOLUME_NONE; ps.volAbs[k].set(0.f); }
  }
}

__device__ __forceinline__ void store_path(const PathSoA& P, uint32_t s, const PathState& ps, bool volumes) {
  P.orgLambda[s] = make_float4(ps.origin.x, ps.origin.y, ps.origin.z, ps.lambda);
  P.dirPdf[s] = make_float4(ps.dir.x, ps.dir.y, ps.dir.z, ps.pdfThisRay);
  P.thr[s] = make_float4(ps.throughput.s[0], ps.throughput.s[1], ps.throughput.s[2], ps.throughput.s[3]);
  P.light[s] = make_float4(ps.totalLight.s[0], ps.totalLight.s[1], ps.totalLight.s[2], ps.totalLight.s[3]);
  P.st0[s] = make_uint4(ps.dim, ps.seed, ps.flags, ps.pathLength);
  P.st1[s] = make_uint4(ps.rayState, ps.nvol, ps.furnaceInstance, 0u);
  if (volumes) {
    P.volIor[s] = make_float4(ps.volIor[0], ps.volIor[1], ps.volIor[2], ps.volIor[3]);
    P.volMat[s] = make_uint4(ps.volMat[0], ps.volMat[1], ps.volMat[2], ps.volMat[3]);
    for (int k = 0; k < 4; ++k)
      P.volAbs[(size_t)s * 4 + k] = make_float4(ps.volAbs[k].s[0], ps.volAbs[k].s[1], ps.volAbs[k].s[2], ps.volAbs[k].s[3]);
  }
}

// Reserve one queue entry per lane with pred set: one atomic per wave, lane offsets from the ballot.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  unsigned long long mask = __ballot(pred ? 1 : 0);
  if (mask == 0ull) return 0u;
  uint32_t lane = __lane_id();
  uint32_t prefix = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, leader);
  return base + prefix;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// Slots [p * npix, (p + 1) * npix) hold the paths of pass p of the batch (several sample indices are
// in flight at once so that every bounce launch carries enough rays to hide traversal latency).
__global__ void k_start(FrameConstants fc, const uint32_t* points, PathSoA P, Queues Q, FilmBuffers F, uint32_t npix,
                        uint32_t nslots, int writeFeatures) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) { Q.counters[0] = nslots; Q.counters[1] = 0; Q.counters[2] = 0; Q.counters[4] = 0; }
  if (i >= nslots) return;
  uint32_t W = (uint32_t)fc.targetTexDimensions[0];
  uint32_t pass = i / npix, pix = i - pass * npix;
  PathState ps;
  start_path(fc, points + (size_t)pass * YPT_SOBOL_POINT_DIMS, pix % W, pix / W, ps);
  store_path(P, i, ps, false);
  Q.ray[0][i] = i;
  if (writeFeatures) {
    FirstHitFeatures none;
    none.hit = 0u;
    store_features(F.features + (size_t)i * FEATURE_STRIDE, none);
  }
}
#endif

template <int WIDTH, bool COUNT>
__global__ void k_extend(DevScene S, PathSoA P, Queues Q, int qsel, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // counters[5]: fetch counter of a persistent shadow launch that may follow (subsurface scenes)
  if (i == 0) { Q.counters[qsel ^ 1] = 0; Q.counters[2] = 0; Q.counters[5] = 0; }
  uint32_t n = Q.counters[qsel];
  if (n < Q.waveBelow) return;  // a short queue: k_extend_w traces it
  TraversalCounters cnt = {0u, 0u};
  if (i < n) {
    uint32_t slot = Q.ray[qsel][i];
    float4 a = P.orgLambda[slot], b = P.dirPdf[slot];
    TravStack st;
    st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
    st.ovfStride = Q.ovfStride; st.gid = i; st.sp = 0;
    float t = 0.f, b1 = 0.f, b2 = 0.f;
    uint32_t tri = NO_HIT_ID;
    vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
    if (WIDTH == 2) traverse_bvh2<0, COUNT>(S.nodes2, S.tris, o, d, RAY_MIN_T, __builtin_inff(), st, t, tri, b1, b2, 0u, cnt);
    else traverse_bvh8<0, COUNT>(S.nodes8, S.tris, o, d, RAY_MIN_T, __builtin_inff(), st, t, tri, b1, b2, 0u, cnt);
    Q.hit[slot] = make_float4(t, b1, b2, __uint_as_float(tri));
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris), nr = wave_sum(i < n ? 1ull : 0ull);
    if (__lane_id() == 0 && (nn | nt)) {
      atomicAdd(&Q.workCounters[0], nn); atomicAdd(&Q.workCounters[1], nt); atomicAdd(&Q.workCounters[4], nr);
    }
  }
}

// A path whose extension ray left the scene: shade_miss and after_segment on the fields they use
// (hero wavelength, direction, throughput, light, path length, ray state); the rest of the slot is
// unchanged, as store_path would write it back.
__device__ __forceinline__ void finish_miss(const SceneView& S, const FrameConstants& fc, const PathSoA& P, uint32_t s) {
  float4 a = P.orgLambda[s], b = P.dirPdf[s], t = P.thr[s], l = P.light[s];
  uint4 u0 = P.st0[s], u1 = P.st1[s];
  PathState ps;
  ps.lambda = a.w;
  ps.dir = vec3(b.x, b.y, b.z);
  ps.throughput = Spectrum(t.x, t.y, t.z, t.w);
  ps.totalLight = Spectrum(l.x, l.y, l.z, l.w);
  ps.pathLength = u0.w;
  ps.rayState = u1.x;
  shade_miss(S, fc, ps);
  ps.pathLength += 1;  // after_segment of a finished path
  P.light[s] = make_float4(ps.totalLight.s[0], ps.totalLight.s[1], ps.totalLight.s[2], ps.totalLight.s[3]);
  P.st0[s] = make_uint4(u0.x, u0.y, u0.z, ps.pathLength);
  P.st1[s] = make_uint4(ps.rayState, u1.y, u1.z, 0u);
}

// Hit compaction for the shading launches. Rays that left the scene cost a few loads and stores,
// and with them in the same wave most lanes of the surface shading would idle. A wave walks the
// queue 64 entries at a time, finishes the misses at once, and parks the slots of hits in LDS until
// 64 of them are waiting; it then shades those 64 with every lane busy. The park buffer holds at
// most 63 + 64 entries. Each workgroup is one wave (64 threads), so the barriers only order its own
// LDS accesses. `take(i)` returns the slot of queue entry i to shade, or NO_SLOT when it has
// finished the entry itself (a miss). next() returns false when the queue is exhausted and nothing
// is parked; otherwise `slot` holds this lane's hit and `act` says whether the lane has one (fewer
// than 64 only at the end).
static constexpr uint32_t NO_SLOT = 0xFFFFFFFFu;
struct HitCompactor {
  uint32_t base, npend, n, stride;
  __device__ __forceinline__ void init(uint32_t count) {
    n = count;
    stride = gridDim.x * blockDim.x;
    base = blockIdx.x * blockDim.x;
    npend = 0;
  }
  template <class Take>
  __device__ __forceinline__ bool next(uint32_t* pend, Take&& take, uint32_t& slot, bool& act) {
    const uint32_t lane = threadIdx.x;
    while (npend < 64 && base < n) {
      uint32_t i = base + lane;
      uint32_t s = i < n ? take(i) : NO_SLOT;
      bool hit = s != NO_SLOT;
      unsigned long long m = __ballot(hit ? 1 : 0);
      if (hit) pend[npend + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = s;
      npend += (uint32_t)__popcll(m);
      base += stride;
    }
    if (npend == 0) return false;
    __syncthreads();
    uint32_t take_n = npend < 64 ? npend : 64;
    npend -= take_n;
    act = lane < take_n;
    slot = act ? pend[npend + lane] : 0u;
    __syncthreads();
    return true;
  }
};

// Grid-stride over the ray queue: a launch sized for the device, not for the slot count, so the
// short queues of deep bounces cost little. The loop bound is uniform across each workgroup, so
// every wave reaches the queue appends with all of its lanes.
// 64-thread workgroups; MINW: waves per SIMD the register allocation must allow
// subsurface probe rays of a shading thread: BVH2 traversal restricted to one instance's triangles
struct GpuProbe {
  static constexpr bool enabled = true;
  const Bvh2Node* nodes2;
  const Bvh8Node* nodes8;
  TriSet tris;
  TravStack st;
  __device__ bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h) {
    float t, b1, b2;
    uint32_t tri;
    bool hit = trace_closest_in_range<2>(nodes2, nodes8, tris, st, o, d, tmin, tmax, first, count, t, tri, b1, b2);
    h.t = t; h.tri = hit ? tri : NO_HIT; h.b1 = b1; h.b2 = b2;
    return hit;
  }
};

// SSS: the scene has subsurface materials. Probe rays then run inside the shading launch, each
// thread with its own stack column in Q.sssStack, and a path may hand two shadow rays (vertex and
// subsurface exit) to the shadow queue, which is sized for that and applied atomically.
// Misses are finished during the walk over the queue (HitCompactor); the loop body shades hits only.
template <bool NEE, int MINW, bool SSS = false>
__global__ __launch_bounds__(64, MINW) void k_shade(FrameConstants fc, DevScene S, const uint32_t* points, uint32_t npix,
                                                    PathSoA P, Queues Q, FilmBuffers F, int qsel, int writeFeatures) {
  __shared__ uint32_t pend[128];
  // first-hit features, when requested, go through LDS rather than registers that would stay
  // allocated across the whole shading step
  __shared__ __attribute__((aligned(16))) unsigned char fhs[64 * sizeof(FirstHitFeatures)];
  FirstHitFeatures* fhp = writeFeatures ? reinterpret_cast<FirstHitFeatures*>(fhs) + threadIdx.x : nullptr;
  if (blockIdx.x == 0 && threadIdx.x == 0) Q.counters[4] = 0;  // fetch counter of the next extension launch
  HitCompactor hc;
  hc.init(Q.counters[qsel]);
  uint32_t slot = 0;
  bool act = false;
  auto take = [&](uint32_t i) {
    uint32_t s = Q.ray[qsel][i];
    if (__float_as_uint(Q.hit[s].w) != NO_HIT_ID) return s;
    finish_miss(S.sv, fc, P, s);
    return NO_SLOT;
  };
  while (hc.next(pend, take, slot, act)) {
    ShadowRay sr, sr2;
    sr.valid = false;
    sr2.valid = false;
    bool cont = false;
    if (act) {
      const uint32_t* point = points + (size_t)(slot / npix) * YPT_SOBOL_POINT_DIMS;
      PathState ps;
      load_path(P, slot, ps);
      uint32_t nvolIn = ps.nvol;
      float4 hv = Q.hit[slot];
      Hit h;
      h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.tri = __float_as_uint(hv.w);
      if (fhp) fhp->hit = 0u;
      bool first = ps.pathLength == 0;
      if constexpr (SSS) {
        GpuProbe pr;
        pr.nodes2 = S.nodes2; pr.nodes8 = S.nodes8; pr.tris = S.tris;
        pr.st.lds = nullptr; pr.st.ovf = Q.sssStack; pr.st.block = 64; pr.st.lane = threadIdx.x; pr.st.ldsDepth = 0;
        pr.st.ovfStride = hc.stride; pr.st.gid = blockIdx.x * blockDim.x + threadIdx.x; pr.st.sp = 0;
        shade_hit(S.sv, fc, point, NEE, ps, h, sr, fhp, &pr, &sr2);
      } else {
        shade_hit(S.sv, fc, point, NEE, ps, h, sr, fhp);
      }
      after_segment(point, ps);
      cont = ps.rayState == RAY_ALIVE && ps.pathLength < fc.maxRayDepth;
      store_path(P, slot, ps, nvolIn > 0 || ps.nvol > 0);
      if (fhp && first && fhp->hit) store_features(F.features + (size_t)slot * FEATURE_STRIDE, *fhp);
    }
    uint32_t si = wave_append(&Q.counters[2], sr.valid);
    if (sr.valid) {
      Q.sOrgTmax[si] = make_float4(sr.origin.x, sr.origin.y, sr.origin.z, sr.tmax);
      Q.sDirTri[si] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, __uint_as_float(sr.targetTri));
      Q.sContrib[si] = make_float4(sr.contrib.s[0], sr.contrib.s[1], sr.contrib.s[2], sr.contrib.s[3]);
      Q.sSlot[si] = slot;
    }
    if constexpr (SSS) {
      uint32_t si2 = wave_append(&Q.counters[2], sr2.valid);
      if (sr2.valid) {
        Q.sOrgTmax[si2] = make_float4(sr2.origin.x, sr2.origin.y, sr2.origin.z, sr2.tmax);
        Q.sDirTri[si2] = make_float4(sr2.dir.x, sr2.dir.y, sr2.dir.z, __uint_as_float(sr2.targetTri));
        Q.sContrib[si2] = make_float4(sr2.contrib.s[0], sr2.contrib.s[1], sr2.contrib.s[2], sr2.contrib.s[3]);
        Q.sSlot[si2] = slot;
      }
    }
    uint32_t ri = wave_append(&Q.counters[qsel ^ 1], cont);
    if (cont) Q.ray[qsel ^ 1][ri] = slot;
  }
}

template <int WIDTH, bool COUNT>
__global__ void k_shadow(DevScene S, PathSoA P, Queues Q, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t n = Q.counters[2];
  if (n < Q.waveBelow) return;  // a short queue: k_shadow_w traces it
  TraversalCounters cnt = {0u, 0u};
  if (i < n) {
    float4 a = Q.sOrgTmax[i], b = Q.sDirTri[i];
    vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
    uint32_t target = __float_as_uint(b.w);
    // the light triangle itself first: the connection counts only if it is the closest hit (a light
    // without geometry, NO_HIT_ID, counts when nothing is hit up to tmax); an alpha-tested light
    // triangle must pass its alpha test where the ray reaches it
    float tt = a.w, b1, b2;
    bool visible = false;
    bool reaches = true;
    if (target != NO_HIT_ID) {
      RayPre rp = ray_precompute(o, d);
      vec3 p1, p2, p3;
      tri_vertices(S.sv, target, p1, p2, p3);
      reaches = intersect_triangle(rp, p1, p2, p3, RAY_MIN_T, a.w, tt, b1, b2) && !alpha_cut(S.sv, target, b1, b2);
    }
    if (reaches) {
      TravStack st;
      st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
      st.ovfStride = Q.ovfStride; st.gid = i; st.sp = 0;
      float bt; uint32_t bid; float c1, c2;
      bool occluded = WIDTH == 2 ? traverse_bvh2<1, COUNT>(S.nodes2, S.tris, o, d, RAY_MIN_T, tt, st, bt, bid, c1, c2, target, cnt)
                                 : traverse_bvh8<1, COUNT>(S.nodes8, S.tris, o, d, RAY_MIN_T, tt, st, bt, bid, c1, c2, target, cnt);
      visible = !occluded;
    }
    if (visible) {
      uint32_t slot = Q.sSlot[i];
      float4 c = Q.sContrib[i];
      float4 l = P.light[slot];
      P.light[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w + c.w);
    }
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris), nr = wave_sum(i < n ? 1ull : 0ull);
    if (__lane_id() == 0 && (nr | nt)) {
      atomicAdd(&Q.workCounters[2], nn); atomicAdd(&Q.workCounters[3], nt); atomicAdd(&Q.workCounters[5], nr);
    }
  }
}

// Persistent traversal: a launch of a few waves per SIMD whose lanes take rays from the queue
// through a fetch counter. A lane whose ray is finished idles until at least `refill` lanes of its
// wave are idle; then the wave takes that many new rays with one atomic. Waves stay busy however
// unevenly long the rays of a bounce are, and a bounce with few rays costs a few hundred waves, not
// one workgroup per path slot. Fetch counter: counters[4] (extension), counters[5] (shadow).
struct WaveFetch {
  uint32_t* counter;
  uint32_t n;
  uint32_t refill;
  uint32_t statics;   // queue entries handed out without the counter: one per lane of the grid
  uint32_t waveBase;  // this wave's first entry among them
  bool started;
  bool drained;
  __device__ __forceinline__ void init(uint32_t* c, uint32_t count, uint32_t refillLanes) {
    counter = c; n = count; refill = refillLanes;
    statics = gridDim.x * blockDim.x;
    waveBase = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    started = false; drained = false;
  }
  // returns the queue index for this lane (>= n when none), or UINT32_MAX when the wave did not fetch.
  // The first call gives every lane the entry of its own global index, so a queue no longer than the
  // grid costs no atomic; entries beyond the grid are taken with one atomic per refill.
  __device__ __forceinline__ uint32_t next(bool idle) {
    if (drained) return 0xFFFFFFFFu;
    uint32_t lane = __lane_id();
    if (!started) {
      started = true;
      if (statics >= n) drained = true;
      return waveBase + lane;
    }
    unsigned long long need = __ballot(idle ? 1 : 0);
    uint32_t k = (uint32_t)__popcll(need);
    if (k < refill) return 0xFFFFFFFFu;
    int leader = __ffsll((long long)need) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, k);
    base = statics + (uint32_t)__shfl((int)base, leader);
    if (base + k >= n) drained = true;
    if (!idle) return 0xFFFFFFFFu;
    return base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
  }
};

template <int WIDTH, bool COUNT, bool ALPHA = true>
__global__ void k_extend_p(DevScene S, PathSoA P, Queues Q, int qsel, uint32_t ldsDepth, uint32_t refill, uint32_t leafShare) {
  extern __shared__ uint32_t lds[];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) { Q.counters[qsel ^ 1] = 0; Q.counters[2] = 0; Q.counters[5] = 0; }
  if (Q.counters[qsel] < Q.waveBelow) return;  // a short queue: k_extend_w traces it
  WaveFetch wf;
  wf.init(&Q.counters[4], Q.counters[qsel], refill);
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = gid; st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  unsigned long long nrays = 0;
  RayTrav r;
  bool has = false;
  uint32_t slot = 0;
  while (true) {
    uint32_t idx = wf.next(!has);
    if (!has && idx < wf.n) {
      slot = Q.ray[qsel][idx];
      float4 a = P.orgLambda[slot], b = P.dirPdf[slot];
      trav_init(r, vec3(a.x, a.y, a.z), vec3(b.x, b.y, b.z), RAY_MIN_T, __builtin_inff(), NO_HIT_ID);
      st.sp = 0;
      has = true;
      if (COUNT) nrays++;
    }
    if (__ballot(has ? 1 : 0) == 0ull) {
      if (wf.drained) break;
      continue;
    }
    // several traversal steps between refill checks: the fetch ballots and the loop's own ballot are
    // scalar work that a single node step does not amortise
    for (uint32_t k = 0; k < Q.stepsPerFetch; ++k) {
      bool done = false;
      if (has) {
        if constexpr (WIDTH == 2) done = step_bvh2_deferred<0, COUNT, ALPHA>(S.nodes2, S.tris, r, st, cnt, leafShare);
        else done = trav_step<WIDTH, 0, COUNT, ALPHA>(S.nodes2, S.nodes8, S.tris, r, st, cnt);
      }
      if (done) {
        Q.hit[slot] = make_float4(r.best, r.b1, r.b2, __uint_as_float(r.bid));
        has = false;
      }
    }
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris), nr = wave_sum(nrays);
    if (__lane_id() == 0 && (nr | nt)) {
      atomicAdd(&Q.workCounters[0], nn); atomicAdd(&Q.workCounters[1], nt); atomicAdd(&Q.workCounters[4], nr);
    }
  }
}

// ATOMIC: two shadow rays of one path slot in one launch (the vertex and the exit point of a
// subsurface event), so the contributions are added atomically. Otherwise each queue entry's slot
// owns its P.light entry in this launch (one ray per path, or one connection record per ray of the
// bidirectional camera wavefront), which takes the plain read-add-write.
template <int WIDTH, bool COUNT, bool ATOMIC = false, bool ALPHA = true>
__global__ void k_shadow_p(DevScene S, PathSoA P, Queues Q, uint32_t ldsDepth, uint32_t refill, uint32_t leafShare) {
  extern __shared__ uint32_t lds[];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  WaveFetch wf;
  if (Q.counters[2] < Q.waveBelow) return;  // a short queue: k_shadow_w traces it
  wf.init(&Q.counters[5], Q.counters[2], refill);
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = gid; st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  unsigned long long nrays = 0;
  RayTrav r;
  bool has = false;
  uint32_t cur = 0;
  while (true) {
    uint32_t idx = wf.next(!has);
    if (!has && idx < wf.n) {
      if (COUNT) nrays++;
      float4 a = Q.sOrgTmax[idx], b = Q.sDirTri[idx];
      vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
      uint32_t target = __float_as_uint(b.w);
      // the light triangle itself first: the connection counts only if it is the closest hit (a light
      // without geometry, NO_HIT_ID, counts when nothing is hit up to tmax)
      float tt = a.w, b1, b2;
      bool reaches = true;
      if (target != NO_HIT_ID) {
        RayPre rp = ray_precompute(o, d);
        vec3 p1, p2, p3;
        tri_vertices(S.sv, target, p1, p2, p3);
        reaches = intersect_triangle(rp, p1, p2, p3, RAY_MIN_T, a.w, tt, b1, b2) && !(ALPHA && alpha_cut(S.sv, target, b1, b2));
      }
      if (reaches) {
        trav_init(r, o, d, RAY_MIN_T, tt, target);
        st.sp = 0;
        has = true;
        cur = idx;
      }
    }
    if (__ballot(has ? 1 : 0) == 0ull) {
      if (wf.drained) break;
      continue;
    }
    // several traversal steps between refill checks: the fetch ballots and the loop's own ballot are
    // scalar work that a single node step does not amortise
    for (uint32_t k = 0; k < Q.stepsPerFetch; ++k) {
      bool done = false;
      if (has) {
        if constexpr (WIDTH == 2) done = step_bvh2_deferred<1, COUNT, ALPHA>(S.nodes2, S.tris, r, st, cnt, leafShare);
        else done = trav_step<WIDTH, 1, COUNT, ALPHA>(S.nodes2, S.nodes8, S.tris, r, st, cnt);
      }
      if (done) {
        if (!r.occluded) {
          uint32_t slot = Q.sSlot[cur];
          float4 c = Q.sContrib[cur];
          if (ATOMIC) {
            float* l = reinterpret_cast<float*>(P.light + slot);
            atomicAdd(l, c.x); atomicAdd(l + 1, c.y); atomicAdd(l + 2, c.z); atomicAdd(l + 3, c.w);
          } else {
            float4 l = P.light[slot];
            P.light[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w + c.w);
          }
        }
        has = false;
      }
    }
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris), nr = wave_sum(nrays);
    if (__lane_id() == 0 && (nr | nt)) {
      atomicAdd(&Q.workCounters[2], nn); atomicAdd(&Q.workCounters[3], nt); atomicAdd(&Q.workCounters[5], nr);
    }
  }
}

// Wave-per-ray forms of the closest-hit and occlusion launches for short queues (fewer entries
// than Q.waveBelow; the thread-per-ray launches of the same queue return at once and these take
// it). One 64-lane workgroup per ray at a time; a wave takes its first ray by its workgroup index
// and the next ones from the queue's fetch counter (counters[4] extension, counters[5] shadow).
// LDS: the lanes' traversal stacks [ldsDepth][64], then the wave's 64-entry node list.
template <int WIDTH, bool COUNT, bool ALPHA = true>
__global__ __launch_bounds__(64) void k_extend_w(DevScene S, PathSoA P, Queues Q, int qsel, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  const uint32_t n = Q.counters[qsel];
  if (n >= Q.waveBelow) return;
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = 64; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = blockIdx.x * 64 + threadIdx.x; st.sp = 0;
  uint32_t* front = lds + ldsDepth * 64;
  const void* nodes = WIDTH == 2 ? (const void*)S.nodes2 : (const void*)S.nodes8;
  TraversalCounters cnt = {0u, 0u};
  unsigned long long nrays = 0;
  uint32_t idx = blockIdx.x;
  while (idx < n) {
    uint32_t slot = Q.ray[qsel][idx];
    float4 a = P.orgLambda[slot], b = P.dirPdf[slot];
    RayTrav r;
    trav_init(r, vec3(a.x, a.y, a.z), vec3(b.x, b.y, b.z), RAY_MIN_T, __builtin_inff(), NO_HIT_ID);
    wave_trace<WIDTH, 0, COUNT, ALPHA>(nodes, S.tris, r, st, front, cnt);
    uint32_t next = 0;
    if (threadIdx.x == 0) {
      Q.hit[slot] = make_float4(r.best, r.b1, r.b2, __uint_as_float(r.bid));
      next = atomicAdd(&Q.counters[4], 1u);
      if (COUNT) nrays++;
    }
    idx = gridDim.x + (uint32_t)__shfl((int)next, 0);
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris);
    if (threadIdx.x == 0 && (nrays | nt)) {
      atomicAdd(&Q.workCounters[0], nn); atomicAdd(&Q.workCounters[1], nt); atomicAdd(&Q.workCounters[4], nrays);
    }
  }
}

template <int WIDTH, bool COUNT, bool ALPHA = true>
__global__ __launch_bounds__(64) void k_shadow_w(DevScene S, PathSoA P, Queues Q, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  const uint32_t n = Q.counters[2];
  if (n >= Q.waveBelow) return;
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = 64; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = blockIdx.x * 64 + threadIdx.x; st.sp = 0;
  uint32_t* front = lds + ldsDepth * 64;
  const void* nodes = WIDTH == 2 ? (const void*)S.nodes2 : (const void*)S.nodes8;
  TraversalCounters cnt = {0u, 0u};
  unsigned long long nrays = 0;
  uint32_t idx = blockIdx.x;
  while (idx < n) {
    float4 a = Q.sOrgTmax[idx], b = Q.sDirTri[idx];
    vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
    uint32_t target = __float_as_uint(b.w);
    // the light triangle itself first, as in k_shadow_p (the same on every lane)
    float tt = a.w, b1, b2;
    bool reaches = true;
    if (target != NO_HIT_ID) {
      RayPre rp = ray_precompute(o, d);
      vec3 p1, p2, p3;
      tri_vertices(S.sv, target, p1, p2, p3);
      reaches = intersect_triangle(rp, p1, p2, p3, RAY_MIN_T, a.w, tt, b1, b2) && !(ALPHA && alpha_cut(S.sv, target, b1, b2));
    }
    bool visible = false;
    if (reaches) {
      RayTrav r;
      trav_init(r, o, d, RAY_MIN_T, tt, target);
      visible = !wave_trace<WIDTH, 1, COUNT, ALPHA>(nodes, S.tris, r, st, front, cnt);
    }
    uint32_t next = 0;
    if (threadIdx.x == 0) {
      if (visible) {
        uint32_t slot = Q.sSlot[idx];
        float4 c = Q.sContrib[idx];
        float4 l = P.light[slot];
        P.light[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w + c.w);
      }
      next = atomicAdd(&Q.counters[5], 1u);
      if (COUNT) nrays++;
    }
    idx = gridDim.x + (uint32_t)__shfl((int)next, 0);
  }
  if (COUNT) {
    unsigned long long nn = wave_sum(cnt.nodes), nt = wave_sum(cnt.tris);
    if (threadIdx.x == 0 && (nrays | nt)) {
      atomicAdd(&Q.workCounters[2], nn); atomicAdd(&Q.workCounters[3], nt); atomicAdd(&Q.workCounters[5], nrays);
    }
  }
}

struct MergeParams {
  uint32_t width, height;
  uint32_t slotBase;          // first path slot of the pass being merged
  uint32_t sampleCount;       // passes merged so far (g_mergeParams.sampleCount)
  uint32_t accumTargets;      // ACCUMULATION_TARGET_COUNT = 2
  float fireflyClampScale;
  float fireflyClampRelaxPower;
  uint32_t killNaN;
  uint32_t cur;               // which merged/count buffer is written this pass
  uint32_t pass;              // pass index within the batch (with a device-side base, see k_merge)
};

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// device-side film bookkeeping of a captured batch: dyn[0] passes merged so far, dyn[1] current buffer
__global__ void k_set_dyn(uint32_t* dyn, uint32_t samples, uint32_t cur) {
  dyn[0] = samples;
  dyn[1] = cur;
}
#endif
#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
__global__ void k_advance_dyn(uint32_t* dyn, uint32_t g) {
  dyn[0] += g;
  dyn[1] ^= (g & 1u);
}
#endif

// clampFireflySample (mergeNewSamples.comp:122-170), reading the previous pass's merged film
__device__ vec3 clamp_firefly(const MergeParams& mp, const FilmBuffers& F, int x, int y, vec3 c, uint32_t accumulated) {
  float scale = mp.fireflyClampScale;
  if (scale <= 0.f || accumulated == 0) return c;
  if (c.x <= 0.f && c.y <= 0.f && c.z <= 0.f) return c;
  const float4* prev = F.merged[mp.cur ^ 1];
  const uint32_t* pc = F.counts[mp.cur ^ 1];
  vec3 ref(0.f);
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      int lx = x + dx, ly = y + dy;
      if (lx < 0 || ly < 0 || lx >= (int)mp.width || ly >= (int)mp.height) continue;
      size_t q = (size_t)ly * mp.width + lx;
      if (pc[q] == 0) continue;
      float4 m = prev[q];
      ref = vmax(ref, vmax(vec3(m.x, m.y, m.z), vec3(0.f)));
    }
  float refLevel = fmax_(fmax_(ref.x, ref.y), ref.z);
  if (refLevel <= 0.f) return c;
  ref = vmax(ref, vec3(refLevel * 0.05f));
  float relax = ypow(fmax_(1.f, (float)accumulated), mp.fireflyClampRelaxPower);
  vec3 allowed = ref * scale * relax;
  float s = 1.f;
  if (c.x > allowed.x) s = fmin_(s, allowed.x / c.x);
  if (c.y > allowed.y) s = fmin_(s, allowed.y / c.y);
  if (c.z > allowed.z) s = fmin_(s, allowed.z / c.z);
  return c * s;
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// path XYZ + mergeNewSamples (mergeNewSamples.comp:173-265) with per-pixel sample counts
// sampleXYZ: XYZ and hero wavelength per texel (bidirectional integrator); null: XYZ of the path slots.
// dyn (captured batches): sample count and current buffer come from the device, advanced by pass.
__global__ void k_merge(DevScene S, PathSoA P, FilmBuffers F, MergeParams mp, const float4* sampleXYZ, const uint32_t* dyn) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= mp.width * mp.height) return;
  if (dyn) {
    mp.sampleCount = dyn[0] + mp.pass;
    mp.cur = dyn[1] ^ (mp.pass & 1u);
  }
  int x = (int)(p % mp.width), y = (int)(p / mp.width);
  vec3 xyz;
  float lam;
  if (sampleXYZ) {
    float4 v = sampleXYZ[p];
    xyz = vec3(v.x, v.y, v.z);
    lam = v.w;
  } else {
    float4 l = P.light[mp.slotBase + p];
    lam = P.orgLambda[mp.slotBase + p].w;
    PathState ps;
    ps.totalLight = Spectrum(l.x, l.y, l.z, l.w);
    ps.lambda = lam;
    xyz = path_xyz(S.sv, ps);
  }
  F.lastSample[p] = make_float4(xyz.x, xyz.y, xyz.z, lam);
  float alpha = 1.f;
  if (mp.killNaN != 0 && (is_nan(xyz.x) || is_nan(xyz.y) || is_nan(xyz.z))) { xyz = vec3(0.f); alpha = 0.f; }
  uint32_t newSamples = (uint32_t)alpha;
  uint32_t oldSamples = F.counts[mp.cur ^ 1][p];
  uint32_t countOut = oldSamples;
  if (newSamples > 0) {
    xyz = clamp_firefly(mp, F, x, y, xyz, oldSamples);
    countOut = newSamples + oldSamples;
    if (F.features && F.featAccum) {  // accumulateFeatures (mergeNewSamples.comp:93-113)
      float4 m[4];
      feature_moments(F.features + (size_t)(mp.slotBase + p) * FEATURE_STRIDE, m);
      float srcBlend = (float)newSamples / (float)(oldSamples + newSamples);
      size_t np = (size_t)mp.width * mp.height;
      for (int k = 0; k < 4; ++k) {
        float4 o = F.featAccum[k * np + p];
        F.featAccum[k * np + p] = oldSamples == 0 ? m[k]
                                                  : make_float4(o.x + srcBlend * (m[k].x - o.x), o.y + srcBlend * (m[k].y - o.y),
                                                                o.z + srcBlend * (m[k].z - o.z), o.w + srcBlend * (m[k].w - o.w));
      }
    }
  }
  F.counts[mp.cur][p] = countOut;
  size_t npix = (size_t)mp.width * mp.height;
  uint32_t layer = mp.sampleCount % mp.accumTargets;
  if (newSamples > 0) {
    uint32_t old = oldSamples / mp.accumTargets;
    float4 dst = F.accum[layer * npix + p];
    double total = (double)(newSamples + old);
    double destBlend = old / total, srcBlend = newSamples / total;
    F.accum[layer * npix + p] = make_float4((float)(dst.x * destBlend + srcBlend * xyz.x), (float)(dst.y * destBlend + srcBlend * xyz.y),
                                            (float)(dst.z * destBlend + srcBlend * xyz.z), 1.f);
  }
  float4 r0 = F.accum[p];
  vec3 sum(r0.x, r0.y, r0.z);
  float cmp = ylog(1.f + sum.y);
  float variance = 0.f;
  uint32_t targets = mp.sampleCount + 1 < mp.accumTargets ? mp.sampleCount + 1 : mp.accumTargets;
  for (uint32_t k = 1; k < targets; ++k) {
    float4 r = F.accum[k * npix + p];
    sum = sum + vec3(r.x, r.y, r.z);
    float lum = ylog(1.f + r.y);
    float d = fabsf(cmp - lum);
    variance += d * d;
  }
  sum = sum / (float)targets;
  variance /= (float)targets;
  F.merged[mp.cur][p] = make_float4(sum.x, sum.y, sum.z, 1.f);
  F.variance[p] = mp.sampleCount == 0 ? 0.f : variance;
}
#endif

// ---------------------------------------------------------------- query kernels used by tests
template <int WIDTH>
__global__ void k_trace_closest(DevScene S, const float* rays, uint32_t n, float* outT, uint32_t* outTri, float* outB,
                                uint32_t* ovf, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* r = rays + (size_t)i * 8;
  TravStack st;
  st.lds = lds; st.ovf = ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = n; st.gid = i; st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  float t = 0.f, b1 = 0.f, b2 = 0.f;
  uint32_t tri = NO_HIT_ID;
  vec3 o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
  if (WIDTH == 2) traverse_bvh2<0, false>(S.nodes2, S.tris, o, d, r[3], r[7], st, t, tri, b1, b2, 0u, cnt);
  else traverse_bvh8<0, false>(S.nodes8, S.tris, o, d, r[3], r[7], st, t, tri, b1, b2, 0u, cnt);
  outT[i] = tri == NO_HIT_ID ? 0.f : t;
  outTri[i] = tri;
  outB[2 * i] = tri == NO_HIT_ID ? 0.f : b1;
  outB[2 * i + 1] = tri == NO_HIT_ID ? 0.f : b2;
}

// the wave-per-ray form of k_trace_closest: workgroups of 64 lanes take rays i = block, block +
// grid, ...; overflow stacks [depth][grid * 64]
template <int WIDTH>
__global__ __launch_bounds__(64) void k_trace_closest_w(DevScene S, const float* rays, uint32_t n, float* outT, uint32_t* outTri,
                                                        float* outB, uint32_t* ovf, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  TravStack st;
  st.lds = lds; st.ovf = ovf; st.block = 64; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = gridDim.x * 64; st.gid = blockIdx.x * 64 + threadIdx.x; st.sp = 0;
  uint32_t* front = lds + ldsDepth * 64;
  const void* nodes = WIDTH == 2 ? (const void*)S.nodes2 : (const void*)S.nodes8;
  TraversalCounters cnt = {0u, 0u};
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const float* r = rays + (size_t)i * 8;
    RayTrav rt;
    trav_init(rt, vec3(r[0], r[1], r[2]), vec3(r[4], r[5], r[6]), r[3], r[7], NO_HIT_ID);
    wave_trace<WIDTH, 0, false, true>(nodes, S.tris, rt, st, front, cnt);
    if (threadIdx.x == 0) {
      bool hit = rt.bid != NO_HIT_ID;
      outT[i] = hit ? rt.best : 0.f;
      outTri[i] = rt.bid;
      outB[2 * i] = hit ? rt.b1 : 0.f;
      outB[2 * i + 1] = hit ? rt.b2 : 0.f;
    }
  }
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
__global__ void k_sample_dims(const uint32_t* point, const uint32_t* seeds, const uint32_t* dims, uint32_t n, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = sample_dim(point, dims[i], seeds[i]);
}
#endif

}  // namespace ypt
