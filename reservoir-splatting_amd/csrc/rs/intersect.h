// Ray/triangle and ray/box tests shared by the GPU traversal and the CPU oracle.
//
// These replace the DXR 1.1 inline RayQuery of the reference
// (Source/Falcor/Scene/RaytracingInline.slang:86-362), whose hardware hit bits cannot be
// reproduced.  The triangle test is the watertight algorithm of Woop, Benthin and Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013, with the double-precision
// fallback for zero edge functions.  A hit is accepted for tMin <= t <= tMax.  Closest
// hits are ordered by (t, triangle id), so the result does not depend on traversal order.
// The box test is the robust slab test of Ize, "Robust BVH Ray Traversal", JCGT 2(2),
// 2013 (far distance scaled by 1 + 2*gamma(3)), conservative for the true geometry; walks
// cull against tMax widened by cull_slack, which also covers the triangle test's own
// rounding of t, so the hits found do not depend on the tree.
#pragma once
#include "common.h"

namespace rs {

struct RayPre {
  vec3 org;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  vec3 invDir;
};

RS_HD float comp(vec3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

RS_HD RayPre ray_pre(vec3 org, vec3 dir) {
  RayPre p;
  vec3 a = vabs(dir);
  int kz = (a.x > a.y) ? ((a.x > a.z) ? 0 : 2) : ((a.y > a.z) ? 1 : 2);
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dz = comp(dir, kz);
  if (dz < 0.f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  p.Sx = comp(dir, kx) / dz;
  p.Sy = comp(dir, ky) / dz;
  p.Sz = 1.f / dz;
  // A zero direction component gets a huge finite inverse instead of inf, so the slab
  // products never form 0 * inf.  Along such an axis the ray's coordinate is constant and
  // the slab test reduces to lo <= o <= hi, which the products get wrong when o lies exactly
  // on a box plane: o == hi gives a far distance of 0 and culls a box whose triangle the
  // watertight test hits on its edge in that plane (a camera's centre row through a shared
  // edge), so the hit found would depend on the tree.  The origin therefore moves off every
  // plane along that axis, by about one float ulp of its magnitude (2^-23 of it, plus 2^-47
  // at zero), for the box and the triangle tests alike: the gap times 1e30 exceeds every
  // distance a walk compares, and a vertex coordinate equal to the moved origin is as rare
  // as any other coincidence.
  float o[3] = {org.x, org.y, org.z};
  const float d[3] = {dir.x, dir.y, dir.z};
  for (int k = 0; k < 3; k++)
    if (d[k] == 0.f) o[k] = o[k] - (fabs_(o[k]) * 0x1p-23f + 0x1p-47f);
  p.org = vec3{o[0], o[1], o[2]};
  p.invDir = vec3{dir.x != 0.f ? 1.f / dir.x : 1e30f, dir.y != 0.f ? 1.f / dir.y : 1e30f,
                  dir.z != 0.f ? 1.f / dir.z : 1e30f};
  return p;
}

// Returns true and writes t, b1, b2 when the ray hits the triangle with tMin <= t <= tMax.
RS_HD bool intersect_triangle(const RayPre& r, vec3 v0, vec3 v1, vec3 v2, float tMin, float tMax, float& t, float& b1,
                              float& b2) {
  vec3 A = v0 - r.org, B = v1 - r.org, C = v2 - r.org;
  float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
  float Ax = comp(A, r.kx) - r.Sx * Akz;
  float Ay = comp(A, r.ky) - r.Sy * Akz;
  float Bx = comp(B, r.kx) - r.Sx * Bkz;
  float By = comp(B, r.ky) - r.Sy * Bkz;
  float Cx = comp(C, r.kx) - r.Sx * Ckz;
  float Cy = comp(C, r.ky) - r.Sy * Ckz;
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
  float rcp = 1.f / det;
  float tt = T * rcp;
  if (!(tt >= tMin && tt <= tMax)) return false;
  t = tt;
  b1 = V * rcp;
  b2 = W * rcp;
  return true;
}

static constexpr float kSlabScale = 1.0000003576278687f;  // 1 + 2*gamma(3), gamma(n) = n*2^-24/(1-n*2^-24)

// Slack added to the far end of the segment that box culling tests (never to the triangle
// test).  The slab test is conservative for the true geometry, but the watertight test's t
// carries an absolute rounding error that grows with the triangle's distance from the ray
// origin (translation, shear and the barycentric sum), not with t: on a segment a few
// micrometres long ending at a surface, it can accept that surface just inside tMax while
// the surface's box starts just beyond, so the hits a walk finds would depend on the tree.
// Culling against tMax + slack, with the slack 2^-14 of the coordinates' magnitude (the
// origin's and the scene's largest), visits every box holding a triangle the test accepts.
RS_HD float cull_slack(vec3 org, float sceneScale) {
  const float o = fmax_(fabs_(org.x), fmax_(fabs_(org.y), fabs_(org.z)));
  return (o + sceneScale) * 0x1p-14f;
}

// Returns the entry distance when the slab test passes, or +inf on a miss.
RS_HD float intersect_box(const RayPre& r, vec3 bmin, vec3 bmax, float tMin, float tMax) {
  float tx0 = (bmin.x - r.org.x) * r.invDir.x, tx1 = (bmax.x - r.org.x) * r.invDir.x;
  float ty0 = (bmin.y - r.org.y) * r.invDir.y, ty1 = (bmax.y - r.org.y) * r.invDir.y;
  float tz0 = (bmin.z - r.org.z) * r.invDir.z, tz1 = (bmax.z - r.org.z) * r.invDir.z;
  float nx = fmin_(tx0, tx1), fx = fmax_(tx0, tx1);
  float ny = fmin_(ty0, ty1), fy = fmax_(ty0, ty1);
  float nz = fmin_(tz0, tz1), fz = fmax_(tz0, tz1);
  float tn = fmax_(nz, fmax_(ny, fmax_(nx, tMin)));
  float tf = fmin_(fz, fmin_(fy, fmin_(fx, tMax)));
  tf = tf * kSlabScale;
  return tn <= tf ? tn : __builtin_inff();
}

}  // namespace rs
