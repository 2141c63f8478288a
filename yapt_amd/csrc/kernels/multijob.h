// Kernels of a fused step: one launch sequence that carries the paths of several render jobs, each
// with its own scene, frame and Sobol points.
//
// The render server gathers the unidirectional (BPT, NEE) jobs that are ready for their next chunk
// of passes and traces them together. Job j owns the path slots [slotBase[j], slotBase[j + 1]) of
// one shared path-state pool: npix_j pixels times the chunk's passes, pass-major as in a job alone.
// The ray and shadow queues hold slots of every job, so a bounce of many small jobs is one launch
// instead of one per job. A lane finds its slot's job from the slot bases (kernel arguments), and
// reads that job's scene, BVH and frame constants from the device job table.
//
// Each path reads only its own job's scene, frame constants and Sobol point, and its own slot, as
// it does when the job renders alone, so a job's film is the same to the bit alone or fused with
// any others (the film merge stays per job, in pass order).
#pragma once
#include "wavefront.h"

namespace ypt {

static constexpr uint32_t FUSED_MAX_JOBS = 16;

// one job of a fused step, in device memory
struct FusedJob {
  DevScene S;
  FrameConstants fc;
  const uint32_t* points;  // Sobol points of the chunk's passes, [pass][YPT_SOBOL_POINT_DIMS]
  uint32_t npix;
  uint32_t nee;            // BPT_WITH_NEE (else BPT)
};

struct FusedArgs {
  const FusedJob* jobs;
  uint32_t njobs;
  uint32_t slotBase[FUSED_MAX_JOBS + 1];  // slotBase[njobs] = slots of the step
  // the job that owns `slot`: a loop of njobs - 1 compares against kernel arguments
  __device__ __forceinline__ uint32_t job_of(uint32_t slot) const {
    uint32_t j = 0;
    for (uint32_t k = 1; k < njobs; ++k) j += slot >= slotBase[k] ? 1u : 0u;
    return j;
  }
};

#ifndef YPT_INSTANCE_TU  // defined once, in the library's host translation unit
// camera rays of every slot of the step (k_start for several jobs)
__global__ void k_start_m(FusedArgs A, PathSoA P, Queues Q) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nslots = A.slotBase[A.njobs];
  if (i == 0) { Q.counters[0] = nslots; Q.counters[1] = 0; Q.counters[2] = 0; Q.counters[4] = 0; }
  if (i >= nslots) return;
  uint32_t j = A.job_of(i);
  const FusedJob& J = A.jobs[j];
  uint32_t local = i - A.slotBase[j];
  uint32_t pass = local / J.npix, pix = local - pass * J.npix;
  uint32_t W = (uint32_t)J.fc.targetTexDimensions[0];
  PathState ps;
  start_path(J.fc, J.points + (size_t)pass * YPT_SOBOL_POINT_DIMS, pix % W, pix / W, ps);
  store_path(P, i, ps, false);
  Q.ray[0][i] = i;
}
#endif

// persistent closest-hit traversal of a queue of several jobs' rays: each lane keeps the BVH of its
// current ray's scene
template <int WIDTH, bool ALPHA>
__global__ void k_extend_pm(FusedArgs A, PathSoA P, Queues Q, int qsel, uint32_t ldsDepth, uint32_t refill, uint32_t leafShare) {
  extern __shared__ uint32_t lds[];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid == 0) { Q.counters[qsel ^ 1] = 0; Q.counters[2] = 0; Q.counters[5] = 0; }
  if (Q.counters[qsel] < Q.waveBelow) return;  // a short queue: k_extend_wm traces it
  WaveFetch wf;
  wf.init(&Q.counters[4], Q.counters[qsel], refill);
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = gid; st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  RayTrav r;
  bool has = false;
  uint32_t slot = 0;
  const Bvh2Node* nodes2 = nullptr;
  const Bvh8Node* nodes8 = nullptr;
  TriSet tris;
  tris.recs = nullptr;
  tris.alpha = AlphaView{nullptr, nullptr, nullptr};
  while (true) {
    uint32_t idx = wf.next(!has);
    if (!has && idx < wf.n) {
      slot = Q.ray[qsel][idx];
      const DevScene& S = A.jobs[A.job_of(slot)].S;
      nodes2 = S.nodes2; nodes8 = S.nodes8; tris = S.tris;
      float4 a = P.orgLambda[slot], b = P.dirPdf[slot];
      trav_init(r, vec3(a.x, a.y, a.z), vec3(b.x, b.y, b.z), RAY_MIN_T, __builtin_inff(), NO_HIT_ID);
      st.sp = 0;
      has = true;
    }
    if (__ballot(has ? 1 : 0) == 0ull) {
      if (wf.drained) break;
      continue;
    }
    for (uint32_t k = 0; k < Q.stepsPerFetch; ++k) {  // node steps between refill checks, as in k_extend_p
      bool done = false;
      if (has) {
        if constexpr (WIDTH == 2) done = step_bvh2_deferred<0, false, ALPHA>(nodes2, tris, r, st, cnt, leafShare);
        else done = trav_step<WIDTH, 0, false, ALPHA>(nodes2, nodes8, tris, r, st, cnt);
      }
      if (done) {
        Q.hit[slot] = make_float4(r.best, r.b1, r.b2, __uint_as_float(r.bid));
        has = false;
      }
    }
  }
}

// shading of a queue of several jobs' paths (k_shade without subsurface probes or first-hit features);
// misses are finished during the walk over the queue, hits shaded in full waves (HitCompactor)
template <int MINW>
__global__ __launch_bounds__(64, MINW) void k_shade_m(FusedArgs A, PathSoA P, Queues Q, int qsel) {
  __shared__ uint32_t pend[128];
  if (blockIdx.x == 0 && threadIdx.x == 0) Q.counters[4] = 0;  // fetch counter of the next extension launch
  HitCompactor hc;
  hc.init(Q.counters[qsel]);
  uint32_t slot = 0;
  bool act = false;
  auto take = [&](uint32_t i) {
    uint32_t s = Q.ray[qsel][i];
    if (__float_as_uint(Q.hit[s].w) != NO_HIT_ID) return s;
    const FusedJob& J = A.jobs[A.job_of(s)];
    finish_miss(J.S.sv, J.fc, P, s);
    return NO_SLOT;
  };
  while (hc.next(pend, take, slot, act)) {
    ShadowRay sr;
    sr.valid = false;
    bool cont = false;
    if (act) {
      uint32_t j = A.job_of(slot);
      const FusedJob& J = A.jobs[j];
      const uint32_t* point = J.points + (size_t)((slot - A.slotBase[j]) / J.npix) * YPT_SOBOL_POINT_DIMS;
      PathState ps;
      load_path(P, slot, ps);
      uint32_t nvolIn = ps.nvol;
      float4 hv = Q.hit[slot];
      Hit h;
      h.t = hv.x; h.b1 = hv.y; h.b2 = hv.z; h.tri = __float_as_uint(hv.w);
      shade_hit(J.S.sv, J.fc, point, J.nee != 0u, ps, h, sr, nullptr);  // fused steps write no features
      after_segment(point, ps);
      cont = ps.rayState == RAY_ALIVE && ps.pathLength < J.fc.maxRayDepth;
      store_path(P, slot, ps, nvolIn > 0 || ps.nvol > 0);
    }
    uint32_t si = wave_append(&Q.counters[2], sr.valid);
    if (sr.valid) {
      Q.sOrgTmax[si] = make_float4(sr.origin.x, sr.origin.y, sr.origin.z, sr.tmax);
      Q.sDirTri[si] = make_float4(sr.dir.x, sr.dir.y, sr.dir.z, __uint_as_float(sr.targetTri));
      Q.sContrib[si] = make_float4(sr.contrib.s[0], sr.contrib.s[1], sr.contrib.s[2], sr.contrib.s[3]);
      Q.sSlot[si] = slot;
    }
    uint32_t ri = wave_append(&Q.counters[qsel ^ 1], cont);
    if (cont) Q.ray[qsel ^ 1][ri] = slot;
  }
}

// persistent occlusion traversal of a queue of several jobs' shadow rays (one per path slot)
template <int WIDTH, bool ALPHA>
__global__ void k_shadow_pm(FusedArgs A, PathSoA P, Queues Q, uint32_t ldsDepth, uint32_t refill, uint32_t leafShare) {
  extern __shared__ uint32_t lds[];
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (Q.counters[2] < Q.waveBelow) return;  // a short queue: k_shadow_wm traces it
  WaveFetch wf;
  wf.init(&Q.counters[5], Q.counters[2], refill);
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = blockDim.x; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = gid; st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  RayTrav r;
  bool has = false;
  uint32_t cur = 0;
  const Bvh2Node* nodes2 = nullptr;
  const Bvh8Node* nodes8 = nullptr;
  TriSet tris;
  tris.recs = nullptr;
  tris.alpha = AlphaView{nullptr, nullptr, nullptr};
  while (true) {
    uint32_t idx = wf.next(!has);
    if (!has && idx < wf.n) {
      const DevScene& S = A.jobs[A.job_of(Q.sSlot[idx])].S;
      float4 a = Q.sOrgTmax[idx], b = Q.sDirTri[idx];
      vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
      uint32_t target = __float_as_uint(b.w);
      // the light triangle itself first, as in k_shadow_p
      float tt = a.w, b1, b2;
      bool reaches = true;
      if (target != NO_HIT_ID) {
        RayPre rp = ray_precompute(o, d);
        vec3 p1, p2, p3;
        tri_vertices(S.sv, target, p1, p2, p3);
        reaches = intersect_triangle(rp, p1, p2, p3, RAY_MIN_T, a.w, tt, b1, b2) && !(ALPHA && alpha_cut(S.sv, target, b1, b2));
      }
      if (reaches) {
        nodes2 = S.nodes2; nodes8 = S.nodes8; tris = S.tris;
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
    for (uint32_t k = 0; k < Q.stepsPerFetch; ++k) {
      bool done = false;
      if (has) {
        if constexpr (WIDTH == 2) done = step_bvh2_deferred<1, false, ALPHA>(nodes2, tris, r, st, cnt, leafShare);
        else done = trav_step<WIDTH, 1, false, ALPHA>(nodes2, nodes8, tris, r, st, cnt);
      }
      if (done) {
        if (!r.occluded) {
          uint32_t slot = Q.sSlot[cur];
          float4 c = Q.sContrib[cur];
          float4 l = P.light[slot];
          P.light[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w + c.w);
        }
        has = false;
      }
    }
  }
}

// wave-per-ray forms for short queues (k_extend_w, k_shadow_w of several jobs): a wave's ray and
// so its scene are the same on every lane
template <int WIDTH, bool ALPHA>
__global__ __launch_bounds__(64) void k_extend_wm(FusedArgs A, PathSoA P, Queues Q, int qsel, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  const uint32_t n = Q.counters[qsel];
  if (n >= Q.waveBelow) return;
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = 64; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = blockIdx.x * 64 + threadIdx.x; st.sp = 0;
  uint32_t* front = lds + ldsDepth * 64;
  TraversalCounters cnt = {0u, 0u};
  uint32_t idx = blockIdx.x;
  while (idx < n) {
    uint32_t slot = Q.ray[qsel][idx];
    const DevScene& S = A.jobs[A.job_of(slot)].S;
    const void* nodes = WIDTH == 2 ? (const void*)S.nodes2 : (const void*)S.nodes8;
    TriSet tris = S.tris;
    float4 a = P.orgLambda[slot], b = P.dirPdf[slot];
    RayTrav r;
    trav_init(r, vec3(a.x, a.y, a.z), vec3(b.x, b.y, b.z), RAY_MIN_T, __builtin_inff(), NO_HIT_ID);
    wave_trace<WIDTH, 0, false, ALPHA>(nodes, tris, r, st, front, cnt);
    uint32_t next = 0;
    if (threadIdx.x == 0) {
      Q.hit[slot] = make_float4(r.best, r.b1, r.b2, __uint_as_float(r.bid));
      next = atomicAdd(&Q.counters[4], 1u);
    }
    idx = gridDim.x + (uint32_t)__shfl((int)next, 0);
  }
}

template <int WIDTH, bool ALPHA>
__global__ __launch_bounds__(64) void k_shadow_wm(FusedArgs A, PathSoA P, Queues Q, uint32_t ldsDepth) {
  extern __shared__ uint32_t lds[];
  const uint32_t n = Q.counters[2];
  if (n >= Q.waveBelow) return;
  TravStack st;
  st.lds = lds; st.ovf = Q.ovf; st.block = 64; st.lane = threadIdx.x; st.ldsDepth = ldsDepth;
  st.ovfStride = Q.ovfStride; st.gid = blockIdx.x * 64 + threadIdx.x; st.sp = 0;
  uint32_t* front = lds + ldsDepth * 64;
  TraversalCounters cnt = {0u, 0u};
  uint32_t idx = blockIdx.x;
  while (idx < n) {
    uint32_t slot = Q.sSlot[idx];
    const DevScene& S = A.jobs[A.job_of(slot)].S;
    float4 a = Q.sOrgTmax[idx], b = Q.sDirTri[idx];
    vec3 o(a.x, a.y, a.z), d(b.x, b.y, b.z);
    uint32_t target = __float_as_uint(b.w);
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
      const void* nodes = WIDTH == 2 ? (const void*)S.nodes2 : (const void*)S.nodes8;
      TriSet tris = S.tris;
      RayTrav r;
      trav_init(r, o, d, RAY_MIN_T, tt, target);
      visible = !wave_trace<WIDTH, 1, false, ALPHA>(nodes, tris, r, st, front, cnt);
    }
    uint32_t next = 0;
    if (threadIdx.x == 0) {
      if (visible) {
        float4 c = Q.sContrib[idx];
        float4 l = P.light[slot];
        P.light[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, l.w + c.w);
      }
      next = atomicAdd(&Q.counters[5], 1u);
    }
    idx = gridDim.x + (uint32_t)__shfl((int)next, 0);
  }
}

}  // namespace ypt
