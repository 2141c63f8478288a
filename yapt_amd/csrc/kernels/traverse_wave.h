// Wave-per-ray traversal: the 64 lanes of a wavefront trace one ray together.
//
// Thread-per-ray traversal (traverse.h) needs many rays in flight to hide the latency of each
// node fetch; the short queues at the end of a bounce sequence, or of a small render job, leave
// most of the device idle while a few long rays finish one dependent fetch at a time. Here one ray
// is spread over a wavefront:
//   1. Frontier expansion. The wave holds a list of nodes whose boxes the ray hits (the root at
//      first). Lane (k, s) = (lane / WIDTH, lane % WIDTH) tests child s of node k, so a BVH2 node
//      takes two lanes and a BVH8 node eight, and each lane intersects the leaf it hits. Inner
//      children that are hit form the next list (ballot and prefix count into LDS). This repeats
//      while the list fits 64 / WIDTH nodes.
//   2. Subtree descent. Once the list is longer (at most 64 nodes), every lane takes one of them
//      and traverses that subtree depth-first with its own stack (the thread-per-ray step
//      functions), so the lanes walk disjoint parts of the tree in parallel.
// Closest hit: before every expansion round and subtree step the lanes agree on the smallest t
// found so far (wave minimum). A lane whose own closest t is larger takes that t as its cut-off
// and forgets its own candidate (id = no hit), so it prunes with the best t of the whole wave but
// still accepts a triangle at exactly that t (a smaller id may win the tie). At the end the wave
// takes the (t, id) minimum over lanes. Every triangle that could be the (t, id) minimum of the
// thread-per-ray traversal is tested under the same inclusive box and triangle cut-offs, so the
// result is the same to the bit.
// Occlusion: the wave stops as soon as any lane finds a triangle other than the target closer
// than the target.
#pragma once
#include "traverse.h"

namespace ypt {

__device__ __forceinline__ float wave_min_f(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}

// the wave's closest t so far becomes every lane's cut-off (closest-hit mode)
__device__ __forceinline__ void wave_share_best(RayTrav& r) {
  float tp = wave_min_f(r.best);
  if (tp < r.best) {
    r.best = tp;
    r.bid = NO_HIT_ID;
  }
}

// (t, id) minimum over the wave, with the barycentrics of the winning lane, on every lane
__device__ __forceinline__ void wave_closest(RayTrav& r) {
  for (int off = 32; off > 0; off >>= 1) {
    float ot = __shfl_xor(r.best, off), o1 = __shfl_xor(r.b1, off), o2 = __shfl_xor(r.b2, off);
    uint32_t oid = (uint32_t)__shfl_xor((int)r.bid, off);
    if (hit_closer(ot, oid, r.best, r.bid)) {
      r.best = ot; r.bid = oid; r.b1 = o1; r.b2 = o2;
    }
  }
}

// One ray (the same RayTrav on every lane of the wave, set up by trav_init) through the tree.
// `front`: 64 words of LDS for this wave's node list; st: the lane's own stack for the subtree
// descent. MODE 0: on return every lane holds the closest hit in r (best, bid, b1, b2). MODE 1:
// returns whether the ray is occluded. Every lane of the wave must call it.
template <int WIDTH, int MODE, bool COUNT, bool ALPHA>
__device__ bool wave_trace(const void* nodes, const TriSet& tris, RayTrav& r, TravStack& st, uint32_t* front,
                           TraversalCounters& cnt) {
  constexpr uint32_t NODE_WORDS = WIDTH == 2 ? 16u : 64u;  // Bvh2Node 64 B, Bvh8Node 256 B
  constexpr uint32_t FAN = 64u / WIDTH;                    // nodes one expansion round takes
  const float* base = reinterpret_cast<const float*>(nodes);
  const uint32_t lane = __lane_id();
  const uint32_t k = lane / WIDTH, s = lane % WIDTH;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t nF = 1;
  if (lane == 0) front[0] = 0u;  // the root
  __syncthreads();
  while (true) {
    uint32_t c = BVH_EMPTY;
    bool hit = false;
    if (k < nF) {
      const float* nf = base + (size_t)front[k] * NODE_WORDS;
      c = reinterpret_cast<const uint32_t*>(nf)[6 * WIDTH + s];
      if (c != BVH_EMPTY) {
        float tn;
        hit = slab_test(vec3(nf[s], nf[2 * WIDTH + s], nf[4 * WIDTH + s]), vec3(nf[WIDTH + s], nf[3 * WIDTH + s], nf[5 * WIDTH + s]),
                        r.o, r.inv, r.tmin, r.best, tn);
      }
      if (COUNT && s == 0) cnt.nodes++;
    }
    bool leaf = hit && (c & BVH_LEAF_BIT) != 0u, inner = hit && !leaf;
    if (leaf) {
      if (MODE == 0) leaf_closest<COUNT, ALPHA>(tris, c, r.rp, r.tmin, r.best, r.bid, r.b1, r.b2, cnt);
      else if (leaf_occluded<COUNT, ALPHA>(tris, c, r.rp, r.tmin, r.best, r.bid, cnt)) r.occluded = 1;
    }
    if (MODE == 1 && __ballot(r.occluded ? 1 : 0) != 0ull) return true;
    unsigned long long m = __ballot(inner ? 1 : 0);
    __syncthreads();  // every lane has read the list
    if (inner) front[__popcll(m & below)] = c;
    __syncthreads();
    if (MODE == 0) wave_share_best(r);
    nF = (uint32_t)__popcll(m);
    if (nF == 0u) break;
    if (nF > FAN) break;
  }
  if (nF > 0u) {
    bool active = lane < nF;
    if (active) {
      r.cur = front[lane];
      st.sp = 0;
    }
    while (__ballot(active ? 1 : 0) != 0ull) {
      if (MODE == 0) wave_share_best(r);
      if (active) {
        bool done;
        if constexpr (WIDTH == 2) done = step_bvh2<MODE, COUNT, ALPHA>(reinterpret_cast<const Bvh2Node*>(nodes), tris, r, st, cnt);
        else done = step_bvh8<MODE, COUNT, ALPHA>(reinterpret_cast<const Bvh8Node*>(nodes), tris, r, st, cnt);
        if (done) active = false;
      }
      if (MODE == 1 && __ballot(r.occluded ? 1 : 0) != 0ull) return true;
    }
  }
  __syncthreads();  // the list is free for the wave's next ray
  if (MODE == 0) wave_closest(r);
  return false;
}

}  // namespace ypt
