// GPU ray queries over a 4-wide BVH: closest hit and any-hit visibility.
//
// Replaces SceneRayQuery::traceRay / traceVisibilityRay of the reference
// (Source/Falcor/Scene/RaytracingInline.slang:86-345, wrappers RS/PathTracer.slang:257-299
// and :1422-1428), which run on DXR RT cores.  CDNA has no ray-tracing units, so the
// traversal is a software loop.  The tree is the GPU-built binary LBVH collapsed to four
// children per node: a node is one 128-byte record holding its children's boxes (SoA per
// axis) and references, so one round of eight 16-byte loads tests four boxes, and a ray
// takes about half the dependent fetches of a binary walk.  Both queries visit the hit
// children nearest first.  The traversal stack lives in LDS (kStackLds entries
// per lane, lane-interleaved so every push/pop is a conflict-free 32-bit LDS access);
// deeper subtrees spill to a private array.  The build checks that the deepest path fits
// the stack (three pushes per level at most).
//
// Ray counting follows logTraceRay semantics (RS/PathTracer.slang:284, :1424): one
// count per closest-hit query and one per visibility query.
#pragma once
#include "../rs/intersect.h"
#include "../rs/camera.h"
#include "../rs/scene.h"

namespace rs {

static constexpr uint32_t kLeafBit = 0x80000000u;

// Binary LBVH node (build stage): both child boxes stored in the parent.
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

// Four-wide node: child k's box is (lo*[k], hi*[k]); child[k] is a node index or a leaf
// code kLeafBit | (first << 3) | (count - 1); slots k >= count are unused.
struct Bvh4Node {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t child[4];
  uint32_t count, _pad[3];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node is two 64-byte lines");
static_assert(sizeof(BvhTri) == 48, "BvhTri is three 16-byte groups");

struct BvhView {
  const Bvh4Node* nodes;
  const BvhTri* tris;  // leaf order
  uint32_t root;       // root reference (node 0, or a leaf code when there is one leaf)
  uint32_t numTris;
  float scale;         // largest vertex coordinate magnitude (cull_slack)
};

#if defined(__HIPCC__)

static constexpr int kBlock = 256;   // threads per block of every tracing kernel
// LDS stack depth (entry and its distance): 24 KiB per 256-lane block, so six blocks fit a CU.
#ifndef RS_STACK_LDS
#define RS_STACK_LDS 12
#endif
static constexpr int kStackLds = RS_STACK_LDS;  // LDS stack entries per lane
static constexpr int kStackPriv = 72;
static constexpr int kStackMax = kStackLds + kStackPriv;

// RS_OUTLINE_TRACE=1 emits each query once per kernel and calls it, instead of inlining
// the walk at every call site.
#if defined(RS_OUTLINE_TRACE) && RS_OUTLINE_TRACE
#define RS_TRACE_FN __noinline__
#else
#define RS_TRACE_FN RS_OUTLINE
#endif

// Closest-hit walks keep each pushed child's entry distance beside it, so an entry that
// lies beyond the closest hit found meanwhile is dropped without fetching its node.

// LDS pointers.  The stack slots are addressed as LDS so that a pop, which reads either an
// LDS slot or the private overflow array, stays two loads of their own (ds_read, scratch)
// instead of one load through a selected generic pointer (flat_load, which also makes the
// compiler wait for every outstanding global load first).
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(3))) float LdsF32;

struct Tracer {
  BvhView bvh;
  LdsU32* lds;       // this lane's first LDS stack slot; slot k is lds[k * kBlock]
  LdsF32* ldsDist;   // entry distances, same layout (closest-hit walks)
  uint32_t rays_closest;
  uint32_t rays_visibility;
};

// Traversal stack: the first kStackLds entries in LDS (lane-interleaved), the rest in a
// private overflow array.  The stack pointer is a plain local of the walk, kept apart from
// the overflow array so that it stays in a register (inside one aggregate with the
// dynamically indexed array it would live in scratch memory).  The build guarantees
// 3 * depth <= kStackMax; the bound check keeps a malformed tree from writing outside the
// arrays.
__device__ __forceinline__ void stack_push(LdsU32* lds, uint32_t* priv, int& sp, uint32_t v) {
  if (sp < kStackLds) lds[sp * kBlock] = v;
  else if (sp < kStackMax) priv[sp - kStackLds] = v;
  else return;
  sp++;
}
__device__ __forceinline__ uint32_t stack_pop(const LdsU32* lds, const uint32_t* priv, int& sp) {
  sp--;
  return sp < kStackLds ? lds[sp * kBlock] : priv[sp - kStackLds];
}
__device__ __forceinline__ void stack_push_dist(LdsU32* lds, LdsF32* ldsDist, uint32_t* priv, float* privDist,
                                                int& sp, uint32_t v, float t) {
  if (sp < kStackLds) {
    lds[sp * kBlock] = v;
    ldsDist[sp * kBlock] = t;
  } else if (sp < kStackMax) {
    priv[sp - kStackLds] = v;
    privDist[sp - kStackLds] = t;
  } else {
    return;
  }
  sp++;
}
__device__ __forceinline__ uint32_t stack_pop_dist(const LdsU32* lds, const LdsF32* ldsDist, const uint32_t* priv,
                                                   const float* privDist, int& sp, float& t) {
  sp--;
  if (sp < kStackLds) {
    t = ldsDist[sp * kBlock];
    return lds[sp * kBlock];
  }
  t = privDist[sp - kStackLds];
  return priv[sp - kStackLds];
}

__device__ __forceinline__ vec3 ld3(const float* p) { return vec3{p[0], p[1], p[2]}; }

// A node's four child boxes as seen by one ray: per axis the plane the ray enters first
// (near) and the one it leaves by (far).  For 1/d >= 0 they are lo and hi, otherwise hi
// and lo, so the walk loads them from the node by the ray's direction signs and the slab
// test needs no per-child min/max.  Rounding is monotonic, so (lo - o) * (1/d) and
// (hi - o) * (1/d) keep the order of lo and hi: the near and far distances are exactly
// the min and max of the two, as in intersect_box.
struct Node4 {
  float4 nx, fx, ny, fy, nz, fz;
  uint4 child;
  uint32_t count;
};

// Byte offsets, within a node record, of each axis's near and far planes for a ray: lo* sit
// at float4 0, 2, 4 and hi* at 1, 3, 5.
struct NearFar {
  uint32_t nx, fx, ny, fy, nz, fz;
};

__device__ __forceinline__ NearFar near_far(vec3 invDir) {
  const uint32_t sx = invDir.x < 0.f ? 16u : 0u, sy = invDir.y < 0.f ? 16u : 0u, sz = invDir.z < 0.f ? 16u : 0u;
  return NearFar{sx, 16u - sx, 32u + sy, 48u - sy, 64u + sz, 80u - sz};
}

// The node's planes, children and count.  Every address is the uniform base of the node
// array plus a 32-bit byte offset, so each load takes the scalar-base form with one 32-bit
// lane offset (the node array stays below 4 GiB; the build checks it).
__device__ __forceinline__ float4 node_f4(const char* base, uint32_t off) {
  return *reinterpret_cast<const float4*>(base + off);
}
// Leaf triangle k, three 16-byte loads from the uniform base of the triangle array (below
// 4 GiB, checked by the build); the id rides in the first one's fourth word.
struct LeafTri {
  vec3 v0, v1, v2;
  uint32_t prim;
};
__device__ __forceinline__ LeafTri load_tri(const BvhTri* tris, uint32_t k) {
  const char* base = reinterpret_cast<const char*>(tris);
  const uint32_t at = k * (uint32_t)sizeof(BvhTri);
  const float4 a = node_f4(base, at), b = node_f4(base, at + 16u), c = node_f4(base, at + 32u);
  return LeafTri{vec3{a.x, a.y, a.z}, vec3{b.x, b.y, b.z}, vec3{c.x, c.y, c.z}, __float_as_uint(a.w)};
}

__device__ __forceinline__ Node4 load_node4(const Bvh4Node* nodes, uint32_t i, const NearFar& o) {
  const char* base = reinterpret_cast<const char*>(nodes);
  const uint32_t at = i * (uint32_t)sizeof(Bvh4Node);
  Node4 n;
  n.nx = node_f4(base, at + o.nx);
  n.fx = node_f4(base, at + o.fx);
  n.ny = node_f4(base, at + o.ny);
  n.fy = node_f4(base, at + o.fy);
  n.nz = node_f4(base, at + o.nz);
  n.fz = node_f4(base, at + o.fz);
  n.child = *reinterpret_cast<const uint4*>(base + (at + 96u));
  n.count = *reinterpret_cast<const uint32_t*>(base + (at + 112u));
  return n;
}

// The slab test of intersect_box (same arithmetic, same conservative far-distance scale)
// on near / far planes, with the hardware min/max, which the compiler fuses into
// v_min3/v_max3.  Box culling only decides what is visited, never which hit is returned,
// so it need not reproduce the CPU comparison chain; the operands are never NaN (finite
// boxes, finite origin, finite inverse direction).
__device__ __forceinline__ float slab_entry(const RayPre& rp, float nx, float ny, float nz, float fx, float fy,
                                            float fz, float tMin, float tMax) {
  float tnx = (nx - rp.org.x) * rp.invDir.x, tfx = (fx - rp.org.x) * rp.invDir.x;
  float tny = (ny - rp.org.y) * rp.invDir.y, tfy = (fy - rp.org.y) * rp.invDir.y;
  float tnz = (nz - rp.org.z) * rp.invDir.z, tfz = (fz - rp.org.z) * rp.invDir.z;
  float tn = __builtin_fmaxf(__builtin_fmaxf(tnx, tny), __builtin_fmaxf(tnz, tMin));
  float tf = __builtin_fminf(__builtin_fminf(tfx, tfy), __builtin_fminf(tfz, tMax));
  return tn <= tf * kSlabScale ? tn : __builtin_inff();
}

__device__ __forceinline__ void box_hits4(const RayPre& rp, const Node4& n, float tMin, float tMax, float t[4]) {
  t[0] = slab_entry(rp, n.nx.x, n.ny.x, n.nz.x, n.fx.x, n.fy.x, n.fz.x, tMin, tMax);
  t[1] = slab_entry(rp, n.nx.y, n.ny.y, n.nz.y, n.fx.y, n.fy.y, n.fz.y, tMin, tMax);
  t[2] = slab_entry(rp, n.nx.z, n.ny.z, n.nz.z, n.fx.z, n.fy.z, n.fz.z, tMin, tMax);
  t[3] = slab_entry(rp, n.nx.w, n.ny.w, n.nz.w, n.fx.w, n.fy.w, n.fz.w, tMin, tMax);
  const float inf = __builtin_inff();
  if (n.count < 4) t[3] = inf;
  if (n.count < 3) t[2] = inf;
}

__device__ __forceinline__ void cswap(float& ta, uint32_t& ca, float& tb, uint32_t& cb) {
  bool s = tb < ta;
  float t = s ? tb : ta;
  tb = s ? ta : tb;
  ta = t;
  uint32_t c = s ? cb : ca;
  cb = s ? ca : cb;
  ca = c;
}

// Next stack entry that may still hold a closer hit (entries whose box entry distance lies
// beyond the scaled culling distance, the closest hit found meanwhile plus the culling
// slack, are dropped); false when none is left.
__device__ __forceinline__ bool pop_live(const Tracer& tr, const uint32_t* priv, const float* privDist, int& sp,
                                         float bestT, uint32_t& cur) {
  while (sp > 0) {
    float te;
    uint32_t c = stack_pop_dist(tr.lds, tr.ldsDist, priv, privDist, sp, te);
    if (te <= bestT * kSlabScale) {
      cur = c;
      return true;
    }
  }
  return false;
}

// Closest hit over [tMin, tMax]; ties in t resolve to the smaller triangle id.
__device__ RS_TRACE_FN HitInfo trace_closest(Tracer& tr, const Ray& ray, float& hitT) {
  tr.rays_closest++;
  HitInfo best = hit_invalid();
  float bestT = ray.tMax;
  if (tr.bvh.numTris == 0) return best;
  RayPre rp = ray_pre(ray.origin, ray.dir);
  const float slack = cull_slack(ray.origin, tr.bvh.scale);
  const NearFar nf = near_far(rp.invDir);
  uint32_t priv[kStackPriv];
  float privDist[kStackPriv];
  int sp = 0;
  const float inf = __builtin_inff();
  uint32_t cur = tr.bvh.root;
  while (true) {
    if (!(cur & kLeafBit)) {
      Node4 n = load_node4(tr.bvh.nodes, cur, nf);
      float t[4];
      box_hits4(rp, n, ray.tMin, bestT + slack, t);
      uint32_t c0 = n.child.x, c1 = n.child.y, c2 = n.child.z, c3 = n.child.w;
      float t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
      // nearest first: sort (t, child) ascending, push the far ones back to front
      cswap(t0, c0, t1, c1);
      cswap(t2, c2, t3, c3);
      cswap(t0, c0, t2, c2);
      cswap(t1, c1, t3, c3);
      cswap(t1, c1, t2, c2);
      if (t0 != inf) {
        if (t3 != inf) stack_push_dist(tr.lds, tr.ldsDist, priv, privDist, sp, c3, t3);
        if (t2 != inf) stack_push_dist(tr.lds, tr.ldsDist, priv, privDist, sp, c2, t2);
        if (t1 != inf) stack_push_dist(tr.lds, tr.ldsDist, priv, privDist, sp, c1, t1);
        cur = c0;
        continue;
      }
    } else {
      uint32_t first = (cur & ~kLeafBit) >> 3, count = (cur & 7u) + 1u;
      for (uint32_t i = 0; i < count; i++) {
        const LeafTri t = load_tri(tr.bvh.tris, first + i);
        float tt, b1, b2;
        if (intersect_triangle(rp, t.v0, t.v1, t.v2, ray.tMin, bestT, tt, b1, b2)) {
          if (tt < bestT || t.prim < best.prim) {
            bestT = tt;
            best = HitInfo{t.prim, b1, b2};
          }
        }
      }
    }
    if (!pop_live(tr, priv, privDist, sp, bestT + slack, cur)) break;
  }
  hitT = bestT;
  return best;
}

// True when nothing blocks the segment [tMin, tMax] (DXR accept-first-hit semantics).
__device__ RS_TRACE_FN bool trace_visible(Tracer& tr, const Ray& ray) {
  tr.rays_visibility++;
  if (tr.bvh.numTris == 0) return true;
  RayPre rp = ray_pre(ray.origin, ray.dir);
  const NearFar nf = near_far(rp.invDir);
  const float cullMax = ray.tMax + cull_slack(ray.origin, tr.bvh.scale);
  uint32_t priv[kStackPriv];
  int sp = 0;
  const float inf = __builtin_inff();
  uint32_t cur = tr.bvh.root;
  while (true) {
    if (!(cur & kLeafBit)) {
      Node4 n = load_node4(tr.bvh.nodes, cur, nf);
      float t[4];
      box_hits4(rp, n, ray.tMin, cullMax, t);
      // nearest first here too: it finds an occluder sooner than the slot order does
      uint32_t c0 = n.child.x, c1 = n.child.y, c2 = n.child.z, c3 = n.child.w;
      float t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
      cswap(t0, c0, t1, c1);
      cswap(t2, c2, t3, c3);
      cswap(t0, c0, t2, c2);
      cswap(t1, c1, t3, c3);
      cswap(t1, c1, t2, c2);
      if (t0 != inf) {
        if (t3 != inf) stack_push(tr.lds, priv, sp, c3);
        if (t2 != inf) stack_push(tr.lds, priv, sp, c2);
        if (t1 != inf) stack_push(tr.lds, priv, sp, c1);
        cur = c0;
        continue;
      }
    } else {
      uint32_t first = (cur & ~kLeafBit) >> 3, count = (cur & 7u) + 1u;
      for (uint32_t i = 0; i < count; i++) {
        const LeafTri t = load_tri(tr.bvh.tris, first + i);
        float tt, b1, b2;
        if (intersect_triangle(rp, t.v0, t.v1, t.v2, ray.tMin, ray.tMax, tt, b1, b2)) return false;
      }
    }
    if (sp == 0) break;
    cur = stack_pop(tr.lds, priv, sp);
  }
  return true;
}

#endif

}  // namespace rs
