// GPU ray queries over the LBVH: closest hit and any-hit visibility.
//
// Replaces SceneRayQuery::traceRay / traceVisibilityRay of the reference
// (Source/Falcor/Scene/RaytracingInline.slang:86-345, wrappers RS/PathTracer.slang:257-299
// and :1422-1428), which run on DXR RT cores.  CDNA has no ray-tracing units, so the
// traversal is a software loop over a binary BVH whose child boxes are stored in the
// parent (one 64-byte node fetch tests both children).  The traversal stack lives in
// LDS (kStackLds entries per lane, lane-interleaved so every push/pop is a conflict-free
// 32-bit LDS access); deeper subtrees spill to a small private array.
//
// Ray counting follows logTraceRay semantics (RS/PathTracer.slang:284, :1424): one
// count per closest-hit query and one per visibility query.
#pragma once
#include "../rs/intersect.h"
#include "../rs/camera.h"
#include "../rs/scene.h"

namespace rs {

static constexpr uint32_t kLeafBit = 0x80000000u;

struct BvhNode {
  float lmin[3], lmax[3];
  float rmin[3], rmax[3];
  uint32_t left, right;  // internal index, or kLeafBit | (first << 3) | (count - 1)
  uint32_t _pad[2];
};

struct BvhTri {
  float v0[3];
  uint32_t prim;
  float v1[3];
  float _p1;
  float v2[3];
  float _p2;
};

struct BvhView {
  const BvhNode* nodes;
  const BvhTri* tris;  // leaf order
  uint32_t root;       // root reference (internal 0 or a leaf code when there is one leaf)
  uint32_t numTris;
};

#if defined(__HIPCC__)

static constexpr int kBlock = 256;   // threads per block of every tracing kernel
static constexpr int kStackLds = 24; // LDS stack entries per lane
static constexpr int kStackPriv = 40;

struct Tracer {
  BvhView bvh;
  uint32_t* lds;  // this lane's first LDS stack slot; slot k is lds[k * kBlock]
  uint32_t rays_closest;
  uint32_t rays_visibility;
};

struct TraceStack {
  uint32_t* lds;
  uint32_t priv[kStackPriv];
  int sp;
  // A binary LBVH over 2^k leaves with 64-bit keys is at most 64 levels deep, so the
  // stack never holds more than kStackLds + kStackPriv entries; the bound check keeps a
  // malformed tree from writing outside the arrays.
  __device__ __forceinline__ void push(uint32_t v) {
    if (sp < kStackLds) lds[sp * kBlock] = v;
    else if (sp < kStackLds + kStackPriv) priv[sp - kStackLds] = v;
    else return;
    sp++;
  }
  __device__ __forceinline__ uint32_t pop() {
    sp--;
    return sp < kStackLds ? lds[sp * kBlock] : priv[sp - kStackLds];
  }
};

__device__ __forceinline__ vec3 ld3(const float* p) { return vec3{p[0], p[1], p[2]}; }

// Closest hit over [tMin, tMax]; ties in t resolve to the smaller triangle id.
__device__ RS_OUTLINE HitInfo trace_closest(Tracer& tr, const Ray& ray, float& hitT) {
  tr.rays_closest++;
  HitInfo best = hit_invalid();
  float bestT = ray.tMax;
  if (tr.bvh.numTris == 0) return best;
  RayPre rp = ray_pre(ray.origin, ray.dir);
  TraceStack st;
  st.lds = tr.lds;
  st.sp = 0;
  uint32_t cur = tr.bvh.root;
  while (true) {
    if (!(cur & kLeafBit)) {
      const BvhNode& n = tr.bvh.nodes[cur];
      float tl = intersect_box(rp, ld3(n.lmin), ld3(n.lmax), ray.tMin, bestT);
      float trr = intersect_box(rp, ld3(n.rmin), ld3(n.rmax), ray.tMin, bestT);
      bool hl = tl != __builtin_inff(), hr = trr != __builtin_inff();
      if (hl && hr) {
        uint32_t nearC = tl <= trr ? n.left : n.right;
        uint32_t farC = tl <= trr ? n.right : n.left;
        st.push(farC);
        cur = nearC;
        continue;
      } else if (hl) {
        cur = n.left;
        continue;
      } else if (hr) {
        cur = n.right;
        continue;
      }
    } else {
      uint32_t first = (cur & ~kLeafBit) >> 3, count = (cur & 7u) + 1u;
      for (uint32_t i = 0; i < count; i++) {
        const BvhTri& t = tr.bvh.tris[first + i];
        float tt, b1, b2;
        if (intersect_triangle(rp, ld3(t.v0), ld3(t.v1), ld3(t.v2), ray.tMin, bestT, tt, b1, b2)) {
          if (tt < bestT || t.prim < best.prim) {
            bestT = tt;
            best = HitInfo{t.prim, b1, b2};
          }
        }
      }
    }
    if (st.sp == 0) break;
    cur = st.pop();
  }
  hitT = bestT;
  return best;
}

// True when nothing blocks the segment [tMin, tMax] (DXR accept-first-hit semantics).
__device__ RS_OUTLINE bool trace_visible(Tracer& tr, const Ray& ray) {
  tr.rays_visibility++;
  if (tr.bvh.numTris == 0) return true;
  RayPre rp = ray_pre(ray.origin, ray.dir);
  TraceStack st;
  st.lds = tr.lds;
  st.sp = 0;
  uint32_t cur = tr.bvh.root;
  while (true) {
    if (!(cur & kLeafBit)) {
      const BvhNode& n = tr.bvh.nodes[cur];
      float tl = intersect_box(rp, ld3(n.lmin), ld3(n.lmax), ray.tMin, ray.tMax);
      float trr = intersect_box(rp, ld3(n.rmin), ld3(n.rmax), ray.tMin, ray.tMax);
      bool hl = tl != __builtin_inff(), hr = trr != __builtin_inff();
      if (hl && hr) {
        st.push(n.right);
        cur = n.left;
        continue;
      } else if (hl) {
        cur = n.left;
        continue;
      } else if (hr) {
        cur = n.right;
        continue;
      }
    } else {
      uint32_t first = (cur & ~kLeafBit) >> 3, count = (cur & 7u) + 1u;
      for (uint32_t i = 0; i < count; i++) {
        const BvhTri& t = tr.bvh.tris[first + i];
        float tt, b1, b2;
        if (intersect_triangle(rp, ld3(t.v0), ld3(t.v1), ld3(t.v2), ray.tMin, ray.tMax, tt, b1, b2)) return false;
      }
    }
    if (st.sp == 0) break;
    cur = st.pop();
  }
  return true;
}

#endif

}  // namespace rs
