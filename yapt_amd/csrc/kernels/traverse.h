// BVH traversal on gfx950: one ray per lane, short traversal stack in LDS with a global-memory
// overflow, both children's boxes fetched in one 64-byte load (BVH2) or eight in 256 bytes (BVH8).
// Leaves are intersected as soon as their box is hit, or, in the persistent launches, together with
// the leaves of the wave's other lanes (step_bvh2_deferred). Closest hit = minimum (t, triangle id); the
// shadow query stops at the first triangle that is closer than the light triangle. A triangle
// whose record carries an alpha word counts only where its albedo alpha passes the cutoff; ALPHA =
// false leaves that test out for scenes without alpha-tested triangles.
#pragma once
#include "../core/intersect.h"
#include "../runtime/bvh.h"

namespace ypt {

// the triangle records of a BVH and what alpha testing of them reads
struct TriSet {
  const TriRecord* recs;
  AlphaView alpha;
};

struct TraversalCounters {
  uint32_t nodes;
  uint32_t tris;
};

struct TravStack {
  uint32_t* lds;      // [depth][block] in LDS
  uint32_t* ovf;      // [depth_ovf][stride] in global memory
  uint32_t block;     // lanes per block
  uint32_t lane;      // threadIdx.x
  uint32_t ldsDepth;
  uint32_t ovfStride; // total threads of the launch
  uint32_t gid;       // global thread id
  uint32_t sp;
  __device__ void push(uint32_t v) {
    if (sp < ldsDepth) lds[sp * block + lane] = v;
    else ovf[(size_t)(sp - ldsDepth) * ovfStride + gid] = v;
    ++sp;
  }
  // A stack with an LDS part (ldsDepth > 0, uniform) reads its LDS entry on every pop, an LDS read
  // rather than a flat access choosing between LDS and global memory; the global overflow is read only
  // for entries beyond the LDS part. A stack without one (lds null, as the subsurface probes use) never
  // touches LDS.
  __device__ uint32_t pop() {
    --sp;
    uint32_t v = 0u;
    if (ldsDepth != 0u) {
      v = lds[(sp < ldsDepth ? sp : ldsDepth - 1u) * block + lane];
      asm volatile("" : "+v"(v));  // a value barrier (no instruction): the two reads are not merged into one flat load
    }
    if (sp >= ldsDepth) v = ovf[(size_t)(sp - ldsDepth) * ovfStride + gid];
    return v;
  }
};

__device__ __forceinline__ TriRecord load_tri(const TriRecord* tris, uint32_t k) {
  const float4* p = reinterpret_cast<const float4*>(tris + k);
  float4 a = p[0], b = p[1], c = p[2];
  TriRecord r;
  r.v0[0] = a.x; r.v0[1] = a.y; r.v0[2] = a.z; r.id = __float_as_uint(a.w);
  r.v1[0] = b.x; r.v1[1] = b.y; r.v1[2] = b.z; r.alpha = __float_as_uint(b.w);
  r.v2[0] = c.x; r.v2[1] = c.y; r.v2[2] = c.z;
  return r;
}

// one triangle record against the closest hit so far
template <bool ALPHA, bool FILTER>
__device__ __forceinline__ void tri_closest(const TriSet& tris, uint32_t k, const RayPre& rp, float tmin, float& best,
                                            uint32_t& bestId, float& bb1, float& bb2, uint32_t lo, uint32_t range) {
  TriRecord r = load_tri(tris.recs, k);
  if (FILTER && r.id - lo >= range) return;  // only triangles [lo, lo + range)
  float t, b1, b2;
  if (intersect_triangle(rp, vec3(r.v0[0], r.v0[1], r.v0[2]), vec3(r.v1[0], r.v1[1], r.v1[2]), vec3(r.v2[0], r.v2[1], r.v2[2]),
                         tmin, best, t, b1, b2) &&
      hit_closer(t, r.id, best, bestId) && (!ALPHA || r.alpha == 0u || alpha_test_passes(tris.alpha, r.alpha - 1u, b1, b2))) {
    best = t; bestId = r.id; bb1 = b1; bb2 = b2;
  }
}

// true when triangle record k is not `target` and is hit closer than (targetT, target)
template <bool ALPHA>
__device__ __forceinline__ bool tri_occludes(const TriSet& tris, uint32_t k, const RayPre& rp, float tmin, float targetT,
                                             uint32_t target) {
  TriRecord r = load_tri(tris.recs, k);
  if (r.id == target) return false;
  float t, b1, b2;
  return intersect_triangle(rp, vec3(r.v0[0], r.v0[1], r.v0[2]), vec3(r.v1[0], r.v1[1], r.v1[2]), vec3(r.v2[0], r.v2[1], r.v2[2]),
                            tmin, targetT, t, b1, b2) &&
         hit_closer(t, r.id, targetT, target) && (!ALPHA || r.alpha == 0u || alpha_test_passes(tris.alpha, r.alpha - 1u, b1, b2));
}

__device__ __forceinline__ uint32_t leaf_first(uint32_t ref) { return ref & 0x07FFFFFFu; }
__device__ __forceinline__ uint32_t leaf_count(uint32_t ref) { return ((ref >> 27) & 15u) + 1u; }

template <bool COUNT, bool ALPHA = true, bool FILTER = false>
__device__ __forceinline__ void leaf_closest(const TriSet& tris, uint32_t ref, const RayPre& rp, float tmin, float& best,
                                             uint32_t& bestId, float& bb1, float& bb2, TraversalCounters& cnt,
                                             uint32_t lo = 0u, uint32_t range = 0u) {
  uint32_t first = leaf_first(ref), count = leaf_count(ref);
  for (uint32_t k = first; k < first + count; ++k) {
    if (COUNT) cnt.tris++;
    tri_closest<ALPHA, FILTER>(tris, k, rp, tmin, best, bestId, bb1, bb2, lo, range);
  }
}

// true when a triangle other than `target` is hit closer than (targetT, target)
template <bool COUNT, bool ALPHA = true>
__device__ __forceinline__ bool leaf_occluded(const TriSet& tris, uint32_t ref, const RayPre& rp, float tmin, float targetT,
                                              uint32_t target, TraversalCounters& cnt) {
  uint32_t first = leaf_first(ref), count = leaf_count(ref);
  for (uint32_t k = first; k < first + count; ++k) {
    if (COUNT) cnt.tris++;
    if (tri_occludes<ALPHA>(tris, k, rp, tmin, targetT, target)) return true;
  }
  return false;
}

// MODE 0: closest hit. MODE 1: occlusion test against (target t, target id).
// Traversal state of one ray; a step visits one node (and intersects the leaves it hits), so a
// lane can hand its ray back and take a new one between steps.
struct RayTrav {
  RayPre rp;
  vec3 o, inv;
  float tmin;
  float best;         // closest t so far (MODE 0) or the target triangle's t (MODE 1)
  float b1, b2;
  uint32_t bid;       // closest triangle so far (MODE 0) or the target triangle (MODE 1)
  uint32_t cur;       // node to visit next
  uint32_t occluded;  // MODE 1 result
  uint32_t lo, range; // FILTER: the closest hit among triangles [lo, lo + range) only
  uint32_t pl0, pl1;  // deferred traversal: hit leaves waiting for the wave's next leaf phase (0: none)
};

__device__ __forceinline__ void trav_init(RayTrav& r, vec3 o, vec3 d, float tmin, float tmax, uint32_t target) {
  r.rp = ray_precompute(o, d);
  r.o = o;
  r.inv = vec3(1.f / d.x, 1.f / d.y, 1.f / d.z);
  r.tmin = tmin;
  r.best = tmax;
  r.b1 = 0.f;
  r.b2 = 0.f;
  r.bid = target;
  r.cur = 0;
  r.occluded = 0;
  r.pl0 = 0;
  r.pl1 = 0;
}

// BVH2: both child boxes in one 64-byte node; hit leaves are intersected at once, the nearer inner
// child is visited next and the farther one pushed. Returns true when the ray is finished.
template <int MODE, bool COUNT, bool ALPHA = true, bool FILTER = false>
__device__ __forceinline__ bool step_bvh2(const Bvh2Node* nodes, const TriSet& tris, RayTrav& r, TravStack& st,
                                          TraversalCounters& cnt) {
  // the node as four 16-byte loads and both boxes tested unconditionally (an empty child's box is
  // ignored), as in step_bvh2_deferred
  const float4* np = reinterpret_cast<const float4*>(nodes + r.cur);
  float4 qx = np[0], qy = np[1], qz = np[2], qc = np[3];
  if (COUNT) cnt.nodes++;
  float tn0, tn1;
  uint32_t c0 = __float_as_uint(qc.x), c1 = __float_as_uint(qc.y);
  bool s0 = slab_test(vec3(qx.x, qy.x, qz.x), vec3(qx.z, qy.z, qz.z), r.o, r.inv, r.tmin, r.best, tn0);
  bool s1 = slab_test(vec3(qx.y, qy.y, qz.y), vec3(qx.w, qy.w, qz.w), r.o, r.inv, r.tmin, r.best, tn1);
  bool h0 = s0 && c0 != BVH_EMPTY, h1 = s1 && c1 != BVH_EMPTY;
  bool l0 = h0 && (c0 & BVH_LEAF_BIT) != 0u, l1 = h1 && (c1 & BVH_LEAF_BIT) != 0u;
  // the triangles of both hit leaves in one loop, so that lanes hitting either leaf share iterations
  uint32_t f0 = leaf_first(c0), n0 = l0 ? leaf_count(c0) : 0u;
  uint32_t f1 = leaf_first(c1), n1 = l1 ? leaf_count(c1) : 0u;
  h0 = h0 && !l0;
  h1 = h1 && !l1;
  for (uint32_t k = 0; k < n0 + n1; ++k) {
    uint32_t rec = k < n0 ? f0 + k : f1 + (k - n0);
    if (COUNT) cnt.tris++;
    if (MODE == 0) tri_closest<ALPHA, FILTER>(tris, rec, r.rp, r.tmin, r.best, r.bid, r.b1, r.b2, r.lo, r.range);
    else if (tri_occludes<ALPHA>(tris, rec, r.rp, r.tmin, r.best, r.bid)) { r.occluded = 1; return true; }
  }
  bool first1 = h1 && (!h0 || tn1 < tn0);   // child 1 is visited next
  if (h0 && h1) st.push(first1 ? c0 : c1);
  if (h0 || h1) {
    r.cur = first1 ? c1 : c0;
    return false;
  }
  if (st.sp == 0) return true;
  r.cur = st.pop();
  return false;
}

// BVH2 with deferred leaves, for the persistent launches, where a wave's lanes run different rays.
// Each step is a node phase or a leaf phase for the whole wave. In a node phase the lanes without
// waiting leaves visit one node; leaves whose boxes are hit are kept in pl0/pl1 and the lane waits.
// A leaf phase comes once at least 1/leafShare of the active lanes wait; they then intersect their
// leaves together, so the triangle loop runs with many lanes instead of the few that hit a leaf in
// one node step. The closest hit is the minimum of (t, id) over the same triangles, so the result
// does not depend on this order. Returns true when the lane's ray is finished.
template <int MODE, bool COUNT, bool ALPHA = true>
__device__ __forceinline__ bool step_bvh2_deferred(const Bvh2Node* nodes, const TriSet& tris, RayTrav& r, TravStack& st,
                                                   TraversalCounters& cnt, uint32_t leafShare) {
  bool wait = (r.pl0 | r.pl1) != 0u;
  uint32_t nWait = (uint32_t)__popcll(__ballot(wait ? 1 : 0)), nActive = (uint32_t)__popcll(__ballot(1));
  if (nWait * leafShare >= nActive) {
    if (!wait) return false;
    uint32_t f0 = leaf_first(r.pl0), n0 = r.pl0 ? leaf_count(r.pl0) : 0u;
    uint32_t f1 = leaf_first(r.pl1), n1 = r.pl1 ? leaf_count(r.pl1) : 0u;
    r.pl0 = 0u;
    r.pl1 = 0u;
    for (uint32_t k = 0; k < n0 + n1; ++k) {
      uint32_t rec = k < n0 ? f0 + k : f1 + (k - n0);
      if (COUNT) cnt.tris++;
      if (MODE == 0) tri_closest<ALPHA, false>(tris, rec, r.rp, r.tmin, r.best, r.bid, r.b1, r.b2, 0u, 0u);
      else if (tri_occludes<ALPHA>(tris, rec, r.rp, r.tmin, r.best, r.bid)) { r.occluded = 1; return true; }
    }
    return r.cur == BVH_EMPTY;
  }
  if (wait) return false;
  // the node as four 16-byte loads, both boxes tested unconditionally (an empty child's box is
  // ignored), and the child choice by selects rather than branches
  const float4* np = reinterpret_cast<const float4*>(nodes + r.cur);
  float4 qx = np[0], qy = np[1], qz = np[2], qc = np[3];
  if (COUNT) cnt.nodes++;
  float tn0, tn1;
  uint32_t c0 = __float_as_uint(qc.x), c1 = __float_as_uint(qc.y);
  bool s0 = slab_test(vec3(qx.x, qy.x, qz.x), vec3(qx.z, qy.z, qz.z), r.o, r.inv, r.tmin, r.best, tn0);
  bool s1 = slab_test(vec3(qx.y, qy.y, qz.y), vec3(qx.w, qy.w, qz.w), r.o, r.inv, r.tmin, r.best, tn1);
  bool h0 = s0 && c0 != BVH_EMPTY, h1 = s1 && c1 != BVH_EMPTY;
  bool l0 = h0 && (c0 & BVH_LEAF_BIT) != 0u, l1 = h1 && (c1 & BVH_LEAF_BIT) != 0u;
  r.pl0 = l0 ? c0 : 0u;
  r.pl1 = l1 ? c1 : 0u;
  h0 = h0 && !l0;
  h1 = h1 && !l1;
  bool both = h0 && h1;
  bool first1 = h1 && (!h0 || tn1 < tn0);   // child 1 is visited next
  uint32_t next = first1 ? c1 : c0;
  if (both) st.push(first1 ? c0 : c1);
  if (!(h0 || h1)) next = st.sp == 0 ? BVH_EMPTY : st.pop();
  r.cur = next;
  return r.cur == BVH_EMPTY && (r.pl0 | r.pl1) == 0u;
}

// BVH8: test all eight boxes of a node, intersect hit leaves, push hit inner children far to near.
template <int MODE, bool COUNT, bool ALPHA = true, bool FILTER = false>
__device__ __forceinline__ bool step_bvh8(const Bvh8Node* nodes, const TriSet& tris, RayTrav& r, TravStack& st,
                                          TraversalCounters& cnt) {
  const float4* np = reinterpret_cast<const float4*>(nodes + r.cur);
  if (COUNT) cnt.nodes++;
  float4 lx0 = np[0], lx1 = np[1], hx0 = np[2], hx1 = np[3];
  float4 ly0 = np[4], ly1 = np[5], hy0 = np[6], hy1 = np[7];
  float4 lz0 = np[8], lz1 = np[9], hz0 = np[10], hz1 = np[11];
  uint4 ch0 = reinterpret_cast<const uint4*>(np)[12], ch1 = reinterpret_cast<const uint4*>(np)[13];
  float lx[8] = {lx0.x, lx0.y, lx0.z, lx0.w, lx1.x, lx1.y, lx1.z, lx1.w};
  float hx[8] = {hx0.x, hx0.y, hx0.z, hx0.w, hx1.x, hx1.y, hx1.z, hx1.w};
  float ly[8] = {ly0.x, ly0.y, ly0.z, ly0.w, ly1.x, ly1.y, ly1.z, ly1.w};
  float hy[8] = {hy0.x, hy0.y, hy0.z, hy0.w, hy1.x, hy1.y, hy1.z, hy1.w};
  float lz[8] = {lz0.x, lz0.y, lz0.z, lz0.w, lz1.x, lz1.y, lz1.z, lz1.w};
  float hz[8] = {hz0.x, hz0.y, hz0.z, hz0.w, hz1.x, hz1.y, hz1.z, hz1.w};
  uint32_t ch[8] = {ch0.x, ch0.y, ch0.z, ch0.w, ch1.x, ch1.y, ch1.z, ch1.w};
  // hit inner children, sorted near to far by insertion (at most 8)
  uint32_t hc[8];
  float ht[8];
  int nh = 0;
  for (int s = 0; s < 8; ++s) {
    if (ch[s] == BVH_EMPTY) continue;
    float tn;
    if (!slab_test(vec3(lx[s], ly[s], lz[s]), vec3(hx[s], hy[s], hz[s]), r.o, r.inv, r.tmin, r.best, tn)) continue;
    if (ch[s] & BVH_LEAF_BIT) {
      if (MODE == 0) leaf_closest<COUNT, ALPHA, FILTER>(tris, ch[s], r.rp, r.tmin, r.best, r.bid, r.b1, r.b2, cnt, r.lo, r.range);
      else if (leaf_occluded<COUNT, ALPHA>(tris, ch[s], r.rp, r.tmin, r.best, r.bid, cnt)) { r.occluded = 1; return true; }
      continue;
    }
    int p = nh++;
    while (p > 0 && ht[p - 1] > tn) { ht[p] = ht[p - 1]; hc[p] = hc[p - 1]; --p; }
    ht[p] = tn; hc[p] = ch[s];
  }
  if (nh > 0) {
    for (int k = nh - 1; k >= 1; --k) st.push(hc[k]);
    r.cur = hc[0];
  } else {
    if (st.sp == 0) return true;
    r.cur = st.pop();
  }
  return false;
}

template <int WIDTH, int MODE, bool COUNT, bool ALPHA = true, bool FILTER = false>
__device__ __forceinline__ bool trav_step(const Bvh2Node* nodes2, const Bvh8Node* nodes8, const TriSet& tris, RayTrav& r,
                                          TravStack& st, TraversalCounters& cnt) {
  if (WIDTH == 2) return step_bvh2<MODE, COUNT, ALPHA, FILTER>(nodes2, tris, r, st, cnt);
  return step_bvh8<MODE, COUNT, ALPHA, FILTER>(nodes8, tris, r, st, cnt);
}

// closest hit among the triangles [first, first + count) (an instance's), for subsurface probes
template <int WIDTH>
__device__ bool trace_closest_in_range(const Bvh2Node* nodes2, const Bvh8Node* nodes8, const TriSet& tris, TravStack& st, vec3 o,
                                 vec3 d, float tmin, float tmax, uint32_t first, uint32_t count, float& t, uint32_t& tri,
                                 float& b1, float& b2) {
  RayTrav r;
  trav_init(r, o, d, tmin, tmax, NO_HIT_ID);
  r.lo = first;
  r.range = count;
  st.sp = 0;
  TraversalCounters cnt = {0u, 0u};
  while (!trav_step<WIDTH, 0, false, true, true>(nodes2, nodes8, tris, r, st, cnt)) {}
  t = r.best; tri = r.bid; b1 = r.b1; b2 = r.b2;
  return r.bid != NO_HIT_ID;
}

// Whole traversal of one ray. MODE 0 returns whether anything was hit (closest in bestT/bestId);
// MODE 1 returns whether a triangle other than `target` lies closer than (tmax, target).
template <int MODE, bool COUNT, bool ALPHA = true>
__device__ bool traverse_bvh2(const Bvh2Node* nodes, TriSet tris, vec3 o, vec3 d, float tmin, float tmax,
                              TravStack& st, float& bestT, uint32_t& bestId, float& b1, float& b2, uint32_t target,
                              TraversalCounters& cnt) {
  RayTrav r;
  trav_init(r, o, d, tmin, tmax, MODE == 0 ? NO_HIT_ID : target);
  st.sp = 0;
  while (!step_bvh2<MODE, COUNT, ALPHA>(nodes, tris, r, st, cnt)) {}
  if (MODE == 1) return r.occluded != 0;
  bestT = r.best; bestId = r.bid; b1 = r.b1; b2 = r.b2;
  return r.bid != NO_HIT_ID;
}

template <int MODE, bool COUNT, bool ALPHA = true>
__device__ bool traverse_bvh8(const Bvh8Node* nodes, TriSet tris, vec3 o, vec3 d, float tmin, float tmax,
                              TravStack& st, float& bestT, uint32_t& bestId, float& b1, float& b2, uint32_t target,
                              TraversalCounters& cnt) {
  RayTrav r;
  trav_init(r, o, d, tmin, tmax, MODE == 0 ? NO_HIT_ID : target);
  st.sp = 0;
  while (!step_bvh8<MODE, COUNT, ALPHA>(nodes, tris, r, st, cnt)) {}
  if (MODE == 1) return r.occluded != 0;
  bestT = r.best; bestId = r.bid; b1 = r.b1; b2 = r.b2;
  return r.bid != NO_HIT_ID;
}

}  // namespace ypt
