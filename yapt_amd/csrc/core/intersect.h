// Ray-triangle and ray-box tests.
//
// The reference traces on driver-built acceleration structures with the hardware intersector
// (AccelerationStructureBuilderDx12.cpp:65-379, TraceRay / RayQuery), whose results cannot be
// reproduced. This build owns the test instead: the watertight ray-triangle test of Woop, Benthin and
// Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013 (shear to ray space, 2-D edge
// functions, double-precision re-evaluation of exactly-zero edge functions, no backface culling).
// It is a pure function of (ray, triangle), so every traversal order finds the same closest hit.
// The closest hit is the minimum of (t, triangle id): ties in t go to the lower id.
// The box test is the robust slab test of Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013, with the
// far distance enlarged by 1 + 2*gamma(3) so that no box containing an accepted hit is culled.
#pragma once
#include "common.h"

namespace ypt {

static constexpr uint32_t NO_HIT_ID = 0xFFFFFFFFu;

struct RayPre {
  vec3 org;
  int kx, ky, kz;
  float Sx, Sy, Sz;
};

YPT_HD RayPre ray_precompute(vec3 o, vec3 d) {
  RayPre r;
  r.org = o;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (d[kz] < 0.f) { int t = kx; kx = ky; ky = t; }
  r.kx = kx; r.ky = ky; r.kz = kz;
  r.Sx = d[kx] / d[kz];
  r.Sy = d[ky] / d[kz];
  r.Sz = 1.0f / d[kz];
  return r;
}

// Hit in (tmin, tmax]: returns t and the weights b1, b2 of v1 and v2.
YPT_HD bool intersect_triangle(const RayPre& r, vec3 v0, vec3 v1, vec3 v2, float tmin, float tmax, float& t, float& b1, float& b2) {
  vec3 A = v0 - r.org, B = v1 - r.org, C = v2 - r.org;
  float Akz = A[r.kz], Bkz = B[r.kz], Ckz = C[r.kz];
  float Ax = A[r.kx] - r.Sx * Akz, Ay = A[r.ky] - r.Sy * Akz;
  float Bx = B[r.kx] - r.Sx * Bkz, By = B[r.ky] - r.Sy * Bkz;
  float Cx = C[r.kx] - r.Sx * Ckz, Cy = C[r.ky] - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.f || V == 0.f || W == 0.f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    U = (float)(CxBy - CyBx);
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    V = (float)(AxCy - AyCx);
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  float det = U + V + W;
  if (det == 0.f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float tt = T / det;
  if (!(tt > tmin && tt <= tmax)) return false;
  t = tt;
  b1 = V / det;
  b2 = W / det;
  return true;
}

// (t, id) lexicographic order of candidate closest hits
YPT_HD bool hit_closer(float t, uint32_t id, float bestT, uint32_t bestId) {
  return t < bestT || (t == bestT && id < bestId);
}

static constexpr float SLAB_FAR_SCALE = 1.000002f;  // well above 1 + 2 * gamma(3), gamma(n) = n*eps/(1-n*eps)

// Slab test: the box [lo, hi] overlaps [tmin, tmax] along the ray. NaN from 0 * inf is ignored by
// fminf/fmaxf, which keeps the test conservative for rays parallel to a slab.
YPT_HD bool slab_test(vec3 lo, vec3 hi, vec3 org, vec3 inv, float tmin, float tmax, float& tnear) {
  float tx0 = (lo.x - org.x) * inv.x, tx1 = (hi.x - org.x) * inv.x;
  float ty0 = (lo.y - org.y) * inv.y, ty1 = (hi.y - org.y) * inv.y;
  float tz0 = (lo.z - org.z) * inv.z, tz1 = (hi.z - org.z) * inv.z;
  float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
  float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
  tnear = tn;
  return tn <= tf * SLAB_FAR_SCALE;
}

}  // namespace ypt
