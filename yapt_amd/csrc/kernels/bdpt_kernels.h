// Kernels of the bidirectional integrator on gfx950. One pass (one Sobol sample index):
//   light    - one light subpath per thread (BDPTTraceLightPath.comp), vertices stored in order
//   headers  - the subpaths that exist, in index order (the reference reserves them with an atomic;
//              a fixed order keeps every result reproducible)
//   camera   - one camera subpath per thread; a 64-thread workgroup covers 64 permuted texels and
//              connects all of them to one light subpath (BDPTTraceCameraPath.comp)
//   splat    - light tracing onto a lens point: one record (texel, XYZ) per light vertex
//              (BDPTCalculatePureLightPaths.comp)
//   sort + accumulate - records sorted by texel with a stable radix sort, then summed per texel in
//              record order and added to the camera sample (BDPTAccumulatePureLightPaths.comp),
//              so the floating-point sum has one fixed order
// Traversal is inline (closest hit, target visibility, occlusion) with the LDS stack of traverse.h.
#pragma once
#include "../core/bdpt.h"
#include "wavefront.h"

namespace ypt {

// COUNT: nodes visited, triangles tested and rays traced, for the algorithmic-byte figures
// SSS: the scene has subsurface materials (probe rays); without, the probe query is compiled out
template <int WIDTH, bool COUNT, bool SSS = false>
struct GpuTracer {
  const Bvh2Node* nodes2;
  const Bvh8Node* nodes8;
  TriSet tris;
  SceneView sv;
  TravStack st;
  TraversalCounters cnt;
  uint32_t rays;
  __device__ bool trace(int mode, vec3 o, vec3 d, float tmin, float tmax, uint32_t target, float& t, uint32_t& id, float& b1,
                        float& b2) {
    if (COUNT) rays++;
    if (mode == 0) {
      if (WIDTH == 2) return traverse_bvh2<0, COUNT>(nodes2, tris, o, d, tmin, tmax, st, t, id, b1, b2, target, cnt);
      return traverse_bvh8<0, COUNT>(nodes8, tris, o, d, tmin, tmax, st, t, id, b1, b2, target, cnt);
    }
    if (WIDTH == 2) return traverse_bvh2<1, COUNT>(nodes2, tris, o, d, tmin, tmax, st, t, id, b1, b2, target, cnt);
    return traverse_bvh8<1, COUNT>(nodes8, tris, o, d, tmin, tmax, st, t, id, b1, b2, target, cnt);
  }
  __device__ void flush(unsigned long long* w) {  // w[0] nodes, w[1] triangles, w[2] rays
    if (!COUNT) return;
    atomicAdd(&w[0], (unsigned long long)cnt.nodes);
    atomicAdd(&w[1], (unsigned long long)cnt.tris);
    atomicAdd(&w[2], (unsigned long long)rays);
  }
  __device__ bool closest(vec3 o, vec3 d, float tmin, float tmax, Hit& h) {
    float t = 0.f, b1 = 0.f, b2 = 0.f;
    uint32_t id = NO_HIT_ID;
    bool hit = trace(0, o, d, tmin, tmax, 0u, t, id, b1, b2);
    h.t = hit ? t : 0.f;
    h.tri = hit ? id : NO_HIT_ID;
    h.b1 = hit ? b1 : 0.f;
    h.b2 = hit ? b2 : 0.f;
    return hit;
  }
  // closest hit is `tri`: `tri` is hit in [tmin, tmax] and nothing orders before it
  __device__ bool hits_target(vec3 o, vec3 d, float tmin, float tmax, uint32_t tri) {
    RayPre rp = ray_precompute(o, d);
    vec3 p1, p2, p3;
    tri_vertices(sv, tri, p1, p2, p3);
    float tt, b1, b2;
    if (!intersect_triangle(rp, p1, p2, p3, tmin, tmax, tt, b1, b2) || alpha_cut(sv, tri, b1, b2)) return false;
    float t;
    uint32_t id;
    return !trace(1, o, d, tmin, tt, tri, t, id, b1, b2);
  }
  __device__ bool unoccluded(vec3 o, vec3 d, float tmin, float tmax) {
    float t, b1, b2;
    uint32_t id;
    return !trace(1, o, d, tmin, tmax, NO_HIT_ID, t, id, b1, b2);
  }
  // subsurface probes: the closest hit among one instance's triangles
  __device__ bool closest_in_range(vec3 o, vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, Hit& h) {
    if constexpr (!SSS) {
      return false;
    } else {
      if (COUNT) rays++;
      float t, b1, b2;
      uint32_t tri;
      bool hit = trace_closest_in_range<WIDTH>(nodes2, nodes8, tris, st, o, d, tmin, tmax, first, count, t, tri, b1, b2);
      h.t = t; h.tri = hit ? tri : NO_HIT; h.b1 = b1; h.b2 = b2;
      return hit;
    }
  }
};

struct BdptBuffers {
  LightNodePacked* nodes;  // [path][BDPT_MAX_LIGHT_VERTICES]
  LightNodeSurface* surf;  // [path][BDPT_MAX_LIGHT_VERTICES]: surface of each vertex after the first
  LightVertexCtx* pre;     // [path][BDPT_MAX_LIGHT_VERTICES]: connection context of each vertex after the
                           // first (LightPrefix over the vertices before it), for k_bdpt_connect
  uint32_t* counts;        // vertices per light subpath (0: none)
  float* lambdas;          // hero wavelength per light subpath
  uint32_t* headers;       // existing subpaths in index order
  uint32_t* headerCount;
  const uint32_t* perm;    // texel permutation (y << 16 | x) per dispatch position
  float4* camXYZ;          // per texel: XYZ of the sample, hero wavelength
  uint32_t* recKeys;       // texel of each splat record (npix: none)
  uint32_t* recVals;       // record index
  float4* recXYZ;
  uint32_t* ovf;
  uint32_t ovfStride;
  uint32_t npaths;
  uint32_t recPerPath;     // light vertices per path - 1
  unsigned long long* work;  // counting launches: [0..2] camera subpaths, [3..5] light subpaths and splats
};

template <int WIDTH, bool COUNT, bool SSS = false>
__device__ __forceinline__ GpuTracer<WIDTH, COUNT, SSS> make_tracer(const DevScene& S, uint32_t* lds, const BdptBuffers& B,
                                                                    uint32_t gid, uint32_t ldsDepth) {
  GpuTracer<WIDTH, COUNT, SSS> tr;
  tr.cnt.nodes = 0;
  tr.cnt.tris = 0;
  tr.rays = 0;
  tr.nodes2 = S.nodes2;
  tr.nodes8 = S.nodes8;
  tr.tris = S.tris;
  tr.sv = S.sv;
  tr.st.lds = lds;
  tr.st.ovf = B.ovf;
  tr.st.block = blockDim.x;
  tr.st.lane = threadIdx.x;
  tr.st.ldsDepth = ldsDepth;
  tr.st.ovfStride = B.ovfStride;
  tr.st.gid = gid;
  tr.st.sp = 0;
  return tr;
}

template <int WIDTH, bool COUNT, bool SSS = false>
__global__ __launch_bounds__(64) void k_bdpt_light(FrameConstants fc, DevScene S, const uint32_t* point, BdptBuffers B, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= B.npaths) return;
  GpuTracer<WIDTH, COUNT, SSS> tr = make_tracer<WIDTH, COUNT, SSS>(S, lds, B, p, ldsDepth);
  float lambda;
  uint32_t c = trace_light_path(S.sv, fc, point, p, tr, B.nodes + (size_t)p * BDPT_MAX_LIGHT_VERTICES, lambda);
  B.counts[p] = c;
  B.lambdas[p] = lambda;
  tr.flush(B.work + 3);
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// one workgroup of 1024 threads: subpaths with vertices, compacted in index order
__global__ void k_bdpt_headers(BdptBuffers B) {
  __shared__ uint32_t sums[1024];
  uint32_t t = threadIdx.x;
  uint32_t chunk = (B.npaths + 1023u) / 1024u;
  uint32_t begin = t * chunk;
  uint32_t end = begin + chunk < B.npaths ? begin + chunk : B.npaths;
  uint32_t c = 0;
  for (uint32_t i = begin; i < end; ++i) c += B.counts[i] > 0u ? 1u : 0u;
  sums[t] = c;
  __syncthreads();
  if (t == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < 1024u; ++k) {
      uint32_t v = sums[k];
      sums[k] = run;
      run += v;
    }
    *B.headerCount = run;
  }
  __syncthreads();
  uint32_t o = sums[t];
  for (uint32_t i = begin; i < end; ++i)
    if (B.counts[i] > 0u) B.headers[o++] = i;
}
#endif

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// surface of every stored light vertex after the first, one thread per (subpath, vertex)
__global__ void k_bdpt_node_surfaces(DevScene S, BdptBuffers B) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t p = i / BDPT_MAX_LIGHT_VERTICES, j = i % BDPT_MAX_LIGHT_VERTICES;
  if (p >= B.npaths || j == 0u || j >= B.counts[p]) return;
  const LightNodePacked* nodes = B.nodes + (size_t)p * BDPT_MAX_LIGHT_VERTICES;
  LightNode n = unpack_node(nodes[j]);
  LightNode prev = unpack_node(nodes[j - 1u]);
  LightNodeSurface out;
  light_node_surface(S.sv, n, prev.pos, spectral_lambdas(B.lambdas[p]), out);
  B.surf[(size_t)p * BDPT_MAX_LIGHT_VERTICES + j] = out;
  if (B.pre) {  // the context the connection loop carries into vertex j, made once per vertex
    LightPrefix pre;
    pre.begin();
    for (uint32_t i = 0; i + 1u < j; ++i) pre.advance(unpack_node(nodes[i]), i);
    pre.advance(prev, j - 1u);
    B.pre[(size_t)p * BDPT_MAX_LIGHT_VERTICES + j] = pre.lv;
  }
}
#endif

// 64 threads per workgroup = one 8x8 dispatch tile
// MINW: waves per SIMD the register allocation must allow (1: up to 512 VGPRs, 2: 256)
template <int WIDTH, bool COUNT, int MINW, bool SSS = false>
__global__ __launch_bounds__(64, MINW) void k_bdpt_camera(FrameConstants fc, DevScene S, const uint32_t* point, BdptBuffers B, float* features,
                              uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1];
  uint32_t gcx = (W + BDPT_TILE - 1u) / BDPT_TILE;
  uint32_t g = blockIdx.x, t = threadIdx.x;
  uint32_t gx = g % gcx, gy = g / gcx;
  uint32_t dx = gx * BDPT_TILE + t % BDPT_TILE, dy = gy * BDPT_TILE + t / BDPT_TILE;
  uint32_t hc = *B.headerCount;
  float lambda;
  const LightNodePacked* nodes = nullptr;
  uint32_t count = 0;
  if (hc > 0) {
    uint32_t p = B.headers[pick_light_path(point, g, hc)];
    lambda = B.lambdas[p];
    nodes = B.nodes + (size_t)p * BDPT_MAX_LIGHT_VERTICES;
    count = B.counts[p];
  } else {
    uint32_t q = B.perm[(gy * BDPT_TILE) * W + gx * BDPT_TILE];
    lambda = fallback_wavelength(fc, point, q & 0xFFFFu, q >> 16);
  }
  if (dx >= W || dy >= H) return;
  uint32_t q = B.perm[dy * W + dx];
  uint32_t px = q & 0xFFFFu, py = q >> 16;
  uint32_t pix = py * W + px;
  GpuTracer<WIDTH, COUNT, SSS> tr = make_tracer<WIDTH, COUNT, SSS>(S, lds, B, g * blockDim.x + t, ldsDepth);
  FirstHitFeatures fh;
  fh.hit = 0u;
  NodeSurfaceTable surfs;
  surfs.p = nodes ? B.surf + (nodes - B.nodes) : nullptr;
  Spectrum L = trace_camera_path(S.sv, fc, point, px, py, lambda, nodes, count, surfs, hc == 0 && t == 0, tr,
                                 features ? &fh : nullptr);
  vec3 xyz = spectrum_to_xyz(L, spectral_lambdas(lambda), S.sv.spec);
  B.camXYZ[pix] = make_float4(xyz.x, xyz.y, xyz.z, lambda);
  tr.flush(B.work);
  if (features) store_features(features + (size_t)pix * FEATURE_STRIDE, fh);
}

// Light tracing (t = 1 splats) with one thread per light vertex: thread t takes vertex j + 1 of the
// light subpath in header slot h (t = h * recPerPath + j) and writes record t. Its MIS prefix is
// advanced over vertices 1 .. j with the updates splat_light_path makes, so the records are the ones
// the per-subpath loop makes; a light subpath's vertices no longer wait on each other's visibility
// rays.
template <int WIDTH, bool COUNT>
__global__ __launch_bounds__(64) void k_bdpt_splat(FrameConstants fc, DevScene S, const uint32_t* point, BdptBuffers B, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.npaths * B.recPerPath) return;
  const uint32_t h = t / B.recPerPath, j = t - h * B.recPerPath;
  const uint32_t npix = (uint32_t)fc.targetTexDimensions[0] * (uint32_t)fc.targetTexDimensions[1];
  const uint32_t hc = *B.headerCount;
  SplatRecord rec;
  rec.texel = npix;
  if (h < hc) {
    const uint32_t p = B.headers[h];
    const uint32_t i = j + 1u;
    if (i < B.counts[p]) {
      const LightNodePacked* nodes = B.nodes + (size_t)p * BDPT_MAX_LIGHT_VERTICES;
      Lambdas L = spectral_lambdas(B.lambdas[p]);
      SplatLens lens;
      splat_lens(fc, point, h, lens);
      LightNode n0 = unpack_node(nodes[0]);
      SplatPrefix pf;
      pf.begin(n0);
      for (uint32_t k = 1; k < i; ++k) pf.advance(unpack_node(nodes[k]));
      LightNode n = unpack_node(nodes[i]);
      GpuTracer<WIDTH, COUNT> tr = make_tracer<WIDTH, COUNT>(S, lds, B, t, ldsDepth);
      NodeSurfaceTable surfs;
      surfs.p = B.surf + (size_t)p * BDPT_MAX_LIGHT_VERTICES;
      splat_vertex(S.sv, fc, hc, L, lens, n0, n, i, pf, surfs, tr, rec);
      tr.flush(B.work + 3);
    }
  }
  const bool valid = rec.texel < npix;
  B.recKeys[t] = valid ? rec.texel : npix;
  B.recVals[t] = t;
  if (valid) B.recXYZ[t] = make_float4(rec.xyz.x, rec.xyz.y, rec.xyz.z, 0.f);
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// sorted records: the first record of each texel sums its run in record order
__global__ void k_bdpt_accum(const uint32_t* keys, const uint32_t* vals, const float4* recXYZ, uint32_t n, uint32_t npix,
                             float4* camXYZ) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t key = keys[k];
  if (key >= npix) return;
  if (k > 0 && keys[k - 1] == key) return;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  for (uint32_t j = k; j < n && keys[j] == key; ++j) {
    float4 r = recXYZ[vals[j]];
    sx += r.x; sy += r.y; sz += r.z;
  }
  if (sx != 0.f || sy != 0.f || sz != 0.f) {
    float4 c = camXYZ[key];
    camXYZ[key] = make_float4(c.x + sx, c.y + sy, c.z + sz, c.w);
  }
}
#endif

// ---------------------------------------------------------------- camera subpaths as a wavefront
// The camera subpaths of a pass as queues in HBM: extension rays go through the persistent
// closest-hit kernel, each vertex is shaded by k_bdpt_cshade (surface, emission, every connection's
// contribution and visibility ray, next direction), and the connection rays go through the
// persistent occlusion kernel. Each connection a vertex queues has a record of its own (zeroed when
// queued; the occlusion kernel adds the contribution when the ray is unoccluded, one ray per record),
// and the next shading launch of the path, or k_bdpt_cfinish, adds the records to the path's total in
// the order they were queued: the order in which the megakernel and the oracle add them, so every
// form of the integrator gives the same bits.
struct CamSoA {
  float4* orgLambda;   // ray origin, hero wavelength
  float4* dirPdf;      // ray direction, pdfForwardCamera
  float4* thr;         // throughput
  float4* total;       // radiance so far
  uint4* st0;          // dim, seed, state flags, vertex index
  uint4* st1;          // alive | previous vertex was a subsurface exit << 1, volumes, light subpath
                       // (0xFFFFFFFF: none), its vertex count
  float4* prev;        // previous vertex position, its forward pdf
  float4* prevN;       // previous vertex normal, pdf ratio sum
  float* deferred;     // deferred ratio sum of the camera subpath
  float4* volIor;
  uint4* volMat;
  float4* volAbs;      // 4 per slot
  float4* conn;        // connection records: connCap per slot
  uint32_t* pending;   // records of the slot's last shaded vertices not yet added to its total
  uint32_t connCap;
  DeferredConnections* defer;  // per slot: the shaded vertex whose connections k_bdpt_connect makes
  uint32_t* connList;          // k_bdpt_connect's work: slot * 16 + light vertex index, length in
                               // Q.counters[6 + qsel] of the shading launch that wrote it
};

// Reserve n entries per lane (n may differ per lane, 0 included): one atomic per wave, the lane's
// first entry from an exclusive scan over the wave.
__device__ __forceinline__ uint32_t wave_append_n(uint32_t* counter, uint32_t n) {
  const uint32_t lane = __lane_id();
  uint32_t incl = n;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t v = (uint32_t)__shfl_up((int)incl, off);
    if (lane >= (uint32_t)off) incl += v;
  }
  const uint32_t total = (uint32_t)__shfl((int)incl, 63);
  uint32_t base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(counter, total);
  base = (uint32_t)__shfl((int)base, 63);
  return base + incl - n;
}

// adds the slot's pending connection records to its total, in the order they were queued
__device__ __forceinline__ void fold_connections(const CamSoA& C, uint32_t s, Spectrum& total) {
  uint32_t n = C.pending[s];
  const float4* r = C.conn + (size_t)s * C.connCap;
  for (uint32_t j = 0; j < n; ++j) {
    float4 v = r[j];
    total = total + Spectrum(v.x, v.y, v.z, v.w);
  }
}

__device__ __forceinline__ void load_cam(const CamSoA& C, uint32_t s, CameraPath& c, uint32_t& lp, uint32_t& count) {
  float4 a = C.orgLambda[s], b = C.dirPdf[s], t = C.thr[s], l = C.total[s], pv = C.prev[s], pn = C.prevN[s];
  uint4 u0 = C.st0[s], u1 = C.st1[s];
  c.rayOrigin = vec3(a.x, a.y, a.z);
  c.lambda = a.w;
  c.rayDir = vec3(b.x, b.y, b.z);
  c.pdfForwardCamera = b.w;
  c.throughput = Spectrum(t.x, t.y, t.z, t.w);
  c.total = Spectrum(l.x, l.y, l.z, l.w);
  c.prevPos = vec3(pv.x, pv.y, pv.z);
  c.prevFwdMIS = pv.w;
  c.prevNrm = vec3(pn.x, pn.y, pn.z);
  c.pdfRatioSum = pn.w;
  c.deferredCam = C.deferred[s];
  c.dim = u0.x; c.seed = u0.y; c.rs.flags = u0.z; c.k = u0.w;
  c.alive = (u1.x & 1u) != 0u;
  c.prevIsExit = (u1.x & 2u) != 0u;
  c.pendingExit = false;  // an exit vertex is shaded in the launch that took the subsurface event
  c.rs.n = u1.y;
  lp = u1.z;
  count = u1.w;
  if (c.rs.n > 0) {
    float4 io = C.volIor[s];
    uint4 vm = C.volMat[s];
    c.rs.ior[0] = io.x; c.rs.ior[1] = io.y; c.rs.ior[2] = io.z; c.rs.ior[3] = io.w;
    c.rs.mat[0] = vm.x; c.rs.mat[1] = vm.y; c.rs.mat[2] = vm.z; c.rs.mat[3] = vm.w;
    for (int k = 0; k < 4; ++k) {
      float4 ab = C.volAbs[(size_t)s * 4 + k];
      c.rs.absorption[k] = Spectrum(ab.x, ab.y, ab.z, ab.w);
    }
  } else {
    for (int k = 0; k < 4; ++k) { c.rs.ior[k] = 1.f; c.rs.mat[k] = VOLUME_NONE; c.rs.absorption[k].set(0.f); }
  }
}

__device__ __forceinline__ void store_cam(const CamSoA& C, uint32_t s, const CameraPath& c, uint32_t lp, uint32_t count,
                                          bool volumes) {
  C.orgLambda[s] = make_float4(c.rayOrigin.x, c.rayOrigin.y, c.rayOrigin.z, c.lambda);
  C.dirPdf[s] = make_float4(c.rayDir.x, c.rayDir.y, c.rayDir.z, c.pdfForwardCamera);
  C.thr[s] = make_float4(c.throughput.s[0], c.throughput.s[1], c.throughput.s[2], c.throughput.s[3]);
  C.total[s] = make_float4(c.total.s[0], c.total.s[1], c.total.s[2], c.total.s[3]);
  C.prev[s] = make_float4(c.prevPos.x, c.prevPos.y, c.prevPos.z, c.prevFwdMIS);
  C.prevN[s] = make_float4(c.prevNrm.x, c.prevNrm.y, c.prevNrm.z, c.pdfRatioSum);
  C.deferred[s] = c.deferredCam;
  C.st0[s] = make_uint4(c.dim, c.seed, c.rs.flags, c.k);
  C.st1[s] = make_uint4((c.alive ? 1u : 0u) | (c.prevIsExit ? 2u : 0u), c.rs.n, lp, count);
  if (volumes) {
    C.volIor[s] = make_float4(c.rs.ior[0], c.rs.ior[1], c.rs.ior[2], c.rs.ior[3]);
    C.volMat[s] = make_uint4(c.rs.mat[0], c.rs.mat[1], c.rs.mat[2], c.rs.mat[3]);
    for (int k = 0; k < 4; ++k)
      C.volAbs[(size_t)s * 4 + k] = make_float4(c.rs.absorption[k].s[0], c.rs.absorption[k].s[1], c.rs.absorption[k].s[2],
                                                c.rs.absorption[k].s[3]);
  }
}

// Slot k of the camera wavefront holds dispatch position (dx, dy). Slots are in tile order: the
// positions of one 8x8 dispatch tile are consecutive (fewer in a partial edge tile), so the lanes of
// a wave mostly connect to one light subpath and read the same light vertices.
__device__ __forceinline__ void slot_position(uint32_t W, uint32_t H, uint32_t k, uint32_t& dx, uint32_t& dy) {
  uint32_t band = BDPT_TILE * W;  // positions in one row of tiles; every row but the last is full
  uint32_t gy = k / band, r = k - gy * band;
  uint32_t th = min((uint32_t)BDPT_TILE, H - gy * BDPT_TILE);
  uint32_t gx = r / (BDPT_TILE * th);
  r -= gx * BDPT_TILE * th;
  uint32_t tw = min((uint32_t)BDPT_TILE, W - gx * BDPT_TILE);
  uint32_t ly = r / tw;
  dx = gx * BDPT_TILE + (r - ly * tw);
  dy = gy * BDPT_TILE + ly;
}

__device__ __forceinline__ uint32_t slot_texel(const FrameConstants& fc, const BdptBuffers& B, uint32_t k) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1];
  uint32_t dx, dy;
  slot_position(W, H, k, dx, dy);
  return B.perm[dy * W + dx];
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// one slot per dispatch position; the 8x8 dispatch tile picks one light subpath
__global__ void k_bdpt_cstart(FrameConstants fc, const uint32_t* point, BdptBuffers B, CamSoA C, Queues Q, float* features) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1];
  uint32_t npix = W * H;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) { Q.counters[0] = npix; Q.counters[1] = 0; Q.counters[2] = 0; Q.counters[4] = 0; Q.counters[6] = 0; Q.counters[7] = 0; }
  if (i >= npix) return;
  uint32_t dx, dy;
  slot_position(W, H, i, dx, dy);
  uint32_t gcx = (W + BDPT_TILE - 1u) / BDPT_TILE;
  uint32_t gx = dx / BDPT_TILE, gy = dy / BDPT_TILE, g = gy * gcx + gx;
  uint32_t t = (dy % BDPT_TILE) * BDPT_TILE + dx % BDPT_TILE;
  uint32_t hc = *B.headerCount;
  float lambda;
  uint32_t lp = 0xFFFFFFFFu, count = 0;
  if (hc > 0) {
    lp = B.headers[pick_light_path(point, g, hc)];
    lambda = B.lambdas[lp];
    count = B.counts[lp];
  } else {
    uint32_t q0 = B.perm[(gy * BDPT_TILE) * W + gx * BDPT_TILE];
    lambda = fallback_wavelength(fc, point, q0 & 0xFFFFu, q0 >> 16);
  }
  uint32_t q = B.perm[dy * W + dx];
  CameraPath c;
  camera_begin(fc, point, q & 0xFFFFu, q >> 16, lambda, hc == 0 && t == 0, c);
  store_cam(C, i, c, lp, count, false);
  C.pending[i] = 0u;
  Q.ray[0][(q >> 16) * W + (q & 0xFFFFu)] = i;  // primary rays queued in screen order: coherent traversal
  if (features) {
    FirstHitFeatures none;
    none.hit = 0u;
    store_features(features + (size_t)((q >> 16) * W + (q & 0xFFFFu)) * FEATURE_STRIDE, none);
  }
}
#endif

// connection rays into the shadow queue (one atomic per wave and connection index), each with a
// zeroed record of the slot that the occlusion kernel fills when the ray is unoccluded
struct QueueSink {
  Queues Q;
  float4* records;  // this slot's records
  uint32_t first;   // index of records[0] among all records
  uint32_t count;
  __device__ __forceinline__ void operator()(CameraPath&, const PendingConnection& pc) {
    uint32_t si = wave_append(&Q.counters[2], true);
    Q.sOrgTmax[si] = make_float4(pc.origin.x, pc.origin.y, pc.origin.z, pc.tmax);
    Q.sDirTri[si] = make_float4(pc.dir.x, pc.dir.y, pc.dir.z, __uint_as_float(pc.tri));
    Q.sContrib[si] = make_float4(pc.contrib.s[0], pc.contrib.s[1], pc.contrib.s[2], pc.contrib.s[3]);
    records[count] = make_float4(0.f, 0.f, 0.f, 0.f);
    Q.sSlot[si] = first + count;
    count++;
  }
};

// One camera vertex of slot `slot`: the previous vertex's connections added, the hit (or miss) shaded,
// the state stored. MISS: the caller knows the ray left the scene (the surface code then drops out).
// Returns whether the path continues; nconn: connections left to k_bdpt_connect (SPLIT).
// fhp: this lane's first-hit feature record (in LDS), or null when no features are written.
template <bool SSS, bool SPLIT, bool MISS>
__device__ __forceinline__ bool cshade_slot(const FrameConstants& fc, const DevScene& S, const uint32_t* point,
                                            const BdptBuffers& B, const CamSoA& C, const Queues& Q, uint32_t slot,
                                            uint32_t stride, float* features, FirstHitFeatures* fhp, uint32_t& nconn) {
  const uint32_t W = (uint32_t)fc.targetTexDimensions[0];
  CameraPath c;
  uint32_t lp, count;
  load_cam(C, slot, c, lp, count);
  fold_connections(C, slot, c.total);  // the previous vertex's visible connections, in queue order
  uint32_t nvolIn = c.rs.n;
  Hit h;
  bool hit = false;
  if constexpr (MISS) {
    h.t = 0.f; h.b1 = 0.f; h.b2 = 0.f; h.tri = NO_HIT_ID;
  } else {
    float4 hv = Q.hit[slot];
    h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.tri = __float_as_uint(hv.w);
    hit = h.tri != NO_HIT_ID;
  }
  const LightNodePacked* nodes = lp != 0xFFFFFFFFu ? B.nodes + (size_t)lp * BDPT_MAX_LIGHT_VERTICES : nullptr;
  QueueSink sink;
  sink.Q = Q;
  sink.first = slot * C.connCap;
  sink.records = C.conn + sink.first;
  sink.count = 0;
  NodeSurfaceTable surfs;
  surfs.p = nodes ? B.surf + (size_t)lp * BDPT_MAX_LIGHT_VERTICES : nullptr;
  if (fhp) fhp->hit = 0u;
  bool first = c.k == 0;
  if constexpr (SSS) {
    GpuProbe pr;
    pr.nodes2 = S.nodes2; pr.nodes8 = S.nodes8; pr.tris = S.tris;
    pr.st.lds = nullptr; pr.st.ovf = Q.sssStack; pr.st.block = 64; pr.st.lane = threadIdx.x; pr.st.ldsDepth = 0;
    pr.st.ovfStride = stride; pr.st.gid = blockIdx.x * blockDim.x + threadIdx.x; pr.st.sp = 0;
    camera_vertex(S.sv, fc, point, c, hit, h, nodes, count, surfs, sink, fhp, pr);
    // the exit point of a subsurface event is the next vertex, with no ray to trace: take it now
    if (c.pendingExit && c.alive) camera_vertex(S.sv, fc, point, c, false, h, nodes, count, surfs, sink, nullptr, pr);
  } else {
    NoSubsurface none;
    if constexpr (SPLIT)
      camera_vertex(S.sv, fc, point, c, hit, h, nodes, count, surfs, sink, fhp, none, ConnectLater{C.defer + slot});
    else
      camera_vertex(S.sv, fc, point, c, hit, h, nodes, count, surfs, sink, fhp, none);
  }
  store_cam(C, slot, c, lp, count, nvolIn > 0 || c.rs.n > 0);
  nconn = 0;
  if constexpr (SPLIT) nconn = C.defer[slot].count;
  C.pending[slot] = SPLIT ? nconn : sink.count;
  if (fhp && first && fhp->hit) {
    uint32_t q = slot_texel(fc, B, slot);
    store_features(features + (size_t)((q >> 16) * W + (q & 0xFFFFu)) * FEATURE_STRIDE, *fhp);
  }
  return c.alive;
}

// Grid-stride over the ray queue, misses finished during the walk and hits shaded in full waves
// (HitCompactor). A miss has no exit vertex to take: an exit is shaded in the launch that took its
// subsurface event, and load_cam starts every launch without one.
// first: the first vertex, whose queue holds every slot; the slots are then shaded in slot (tile)
// order rather than the screen order of the primary-ray queue
// SSS: the scene has subsurface materials; probes then run here (stacks in Q.sssStack) and the exit
// vertex of a subsurface event is shaded in the same launch, so a path can queue the connections of
// two vertices
// SPLIT: the vertex's connections are left to k_bdpt_connect (its context goes to C.defer[slot]).
template <int MINW, bool SSS = false, bool SPLIT = false>
__global__ __launch_bounds__(64, MINW) void k_bdpt_cshade(FrameConstants fc, DevScene S, const uint32_t* point, BdptBuffers B,
                                                          CamSoA C, Queues Q, int qsel, int first, float* features) {
  __shared__ uint32_t pend[128];
  // first-hit features, when requested, through LDS (no registers held for them across the step)
  __shared__ __attribute__((aligned(16))) unsigned char fhs[64 * sizeof(FirstHitFeatures)];
  FirstHitFeatures* fhp = features ? reinterpret_cast<FirstHitFeatures*>(fhs) + threadIdx.x : nullptr;
  // counters[6 + qsel]: connection list of this launch (zero since the launch before the last one);
  // counters[6 + (qsel ^ 1)]: the next launch's, read by the connection launch before this one
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Q.counters[4] = 0;
    if constexpr (SPLIT) Q.counters[6 + (qsel ^ 1)] = 0;
  }
  HitCompactor hc;
  hc.init(Q.counters[qsel]);
  auto take = [&](uint32_t i) {
    uint32_t s = first ? i : Q.ray[qsel][i];
    if (__float_as_uint(Q.hit[s].w) != NO_HIT_ID) return s;
    uint32_t nc;
    cshade_slot<SSS, SPLIT, true>(fc, S, point, B, C, Q, s, hc.stride, features, fhp, nc);
    return NO_SLOT;
  };
  uint32_t slot = 0;
  bool act = false;
  while (hc.next(pend, take, slot, act)) {
    bool cont = false;
    uint32_t nconn = 0;
    if (act) cont = cshade_slot<SSS, SPLIT, false>(fc, S, point, B, C, Q, slot, hc.stride, features, fhp, nconn);
    uint32_t ri = wave_append(&Q.counters[qsel ^ 1], cont);
    if (cont) Q.ray[qsel ^ 1][ri] = slot;
    if constexpr (SPLIT) {
      uint32_t ci = wave_append_n(&Q.counters[6 + qsel], nconn);
      for (uint32_t j = 0; j < nconn; ++j) C.connList[ci + j] = slot * BDPT_MAX_LIGHT_VERTICES + j;
    }
  }
}

// The connections of the vertices that k_bdpt_cshade<..., SPLIT> shaded: one thread per entry of
// its list, (slot, light vertex j) for every j below the light subpath's vertex count, so the
// connections of one camera vertex run side by side instead of in one thread's loop. Thread j reads
// the light-vertex context that the loop would carry into vertex j (k_bdpt_node_surfaces advances it
// over vertices 0 .. j-1 once per light vertex, with the loop's own steps) and makes connection j
// with the loop's own step (connect_camera_to_node). Connection j writes record j of the slot (zero, filled by the occlusion
// kernel when unoccluded); the next shading launch adds records 0 .. count-1 in order, the same sum
// as the loop's records with zero terms between them (the totals are never -0, so a +0 term changes
// no bit).
template <int MINW>
__global__ __launch_bounds__(64, MINW) void k_bdpt_connect(FrameConstants fc, DevScene S, BdptBuffers B, CamSoA C, Queues Q, int qsel) {
  constexpr uint32_t PER = BDPT_MAX_LIGHT_VERTICES;
  const uint32_t n = Q.counters[6 + qsel];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t t = base + threadIdx.x;
    bool valid = false;
    PendingConnection pc;
    uint32_t rec = 0;
    if (t < n) {
      const uint32_t v = C.connList[t];
      const uint32_t slot = v / PER, j = v % PER;
      rec = slot * C.connCap + j;
      C.conn[rec] = make_float4(0.f, 0.f, 0.f, 0.f);
      const DeferredConnections& d = C.defer[slot];
      const uint32_t lp = C.st1[slot].z;
      const LightNodePacked* nodes = B.nodes + (size_t)lp * PER;
      NodeSurfaceTable surfs;
      surfs.p = B.surf + (size_t)lp * PER;
      Lambdas L = spectral_lambdas(C.orgLambda[slot].w);
      LightPrefix pre;
      pre.begin();
      if (j > 0) pre.lv = B.pre[(size_t)lp * PER + j];
      LightNode nd = unpack_node(nodes[j]);
      valid = connect_camera_to_node(S.sv, fc, d.k, d.cv, d.ctx, d.rs, L, nd, j, pre.lv, surfs, pc);
    }
    uint32_t si = wave_append(&Q.counters[2], valid);
    if (valid) {
      Q.sOrgTmax[si] = make_float4(pc.origin.x, pc.origin.y, pc.origin.z, pc.tmax);
      Q.sDirTri[si] = make_float4(pc.dir.x, pc.dir.y, pc.dir.z, __uint_as_float(pc.tri));
      Q.sContrib[si] = make_float4(pc.contrib.s[0], pc.contrib.s[1], pc.contrib.s[2], pc.contrib.s[3]);
      Q.sSlot[si] = rec;
    }
  }
}

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// XYZ of every camera subpath into its texel
__global__ void k_bdpt_cfinish(FrameConstants fc, DevScene S, BdptBuffers B, CamSoA C) {
  uint32_t W = (uint32_t)fc.targetTexDimensions[0], H = (uint32_t)fc.targetTexDimensions[1];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  float4 l = C.total[i];
  Spectrum total(l.x, l.y, l.z, l.w);
  fold_connections(C, i, total);  // the last shaded vertex's visible connections
  float lambda = C.orgLambda[i].w;
  vec3 xyz = spectrum_to_xyz(total, spectral_lambdas(lambda), S.sv.spec);
  uint32_t q = slot_texel(fc, B, i);
  B.camXYZ[(q >> 16) * W + (q & 0xFFFFu)] = make_float4(xyz.x, xyz.y, xyz.z, lambda);
}
#endif

}  // namespace ypt
