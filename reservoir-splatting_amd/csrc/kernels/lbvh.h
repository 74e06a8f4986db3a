// GPU-built linear BVH (Karras 2012, "Maximizing Parallelism in the Construction of
// BVHs, Octrees, and k-d Trees") with an atomic bottom-up box pass.
//
// Build: 30-bit Morton codes of triangle centroids in the scene box -> stable radix sort
// of (code, triangle id) -> one Karras hierarchy thread per internal node (keys made
// unique by appending the triangle id) -> bottom-up refit that writes both child boxes
// into the parent (the second child to arrive computes the parent) -> child references
// of subtrees with <= kLeafMax triangles become leaf ranges.  Min/max are exact, so
// the tree does not depend on thread scheduling.
#pragma once
#include "../pass/trace.h"

namespace rs {

// Largest leaf (triangles per leaf range); at most 8 by the leaf-code format.  Two is the
// fastest at 1080p on the 1M-triangle field (1: 47.7, 2: 50.5, 3: 48.9, 4: 46.2, 6: 42.5 fps).
#ifndef RS_LEAF_MAX
#define RS_LEAF_MAX 2
#endif
static constexpr uint32_t kLeafMax = RS_LEAF_MAX;
static_assert(kLeafMax >= 1 && kLeafMax <= 8, "leaf codes hold counts 1..8");

struct LbvhBuild {
  uint32_t n;
  const float* positions;
  const uint32_t* indices;
  vec3 sceneMin, sceneExtentInv;
  uint32_t* morton;     // n
  uint32_t* order;      // n (triangle id in sorted order)
  uint32_t* child;      // 2 * (n - 1): raw children, leaf k encoded as kLeafBit | k
  uint32_t* rangeLo;    // n - 1
  uint32_t* rangeHi;    // n - 1
  uint32_t* parent;     // (n - 1) internal + n leaves: parent of internal i at [i], of leaf k at [n - 1 + k]
  uint32_t* visits;     // n - 1
  float* nodeBox;       // 6 * (n - 1)
  BvhNode* nodes;       // n - 1
  BvhTri* tris;         // n
};

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ void tri_verts(const LbvhBuild& b, uint32_t t, vec3& a, vec3& c, vec3& d) {
  uint32_t i0 = b.indices[3 * t], i1 = b.indices[3 * t + 1], i2 = b.indices[3 * t + 2];
  a = load3(b.positions, i0);
  c = load3(b.positions, i1);
  d = load3(b.positions, i2);
}

__global__ void k_lbvh_morton(LbvhBuild b) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.n) return;
  vec3 a, c, d;
  tri_verts(b, t, a, c, d);
  vec3 lo = vmin(a, vmin(c, d)), hi = vmax(a, vmax(c, d));
  vec3 cen = (lo + hi) * 0.5f;
  vec3 u = (cen - b.sceneMin) * b.sceneExtentInv;
  uint32_t x = (uint32_t)fmin_(fmax_(u.x * 1024.f, 0.f), 1023.f);
  uint32_t y = (uint32_t)fmin_(fmax_(u.y * 1024.f, 0.f), 1023.f);
  uint32_t z = (uint32_t)fmin_(fmax_(u.z * 1024.f, 0.f), 1023.f);
  b.morton[t] = (expand_bits10(x) << 2) | (expand_bits10(y) << 1) | expand_bits10(z);
  b.order[t] = t;
}

__device__ __forceinline__ int lbvh_delta(const LbvhBuild& b, int i, int j) {
  if (j < 0 || j >= (int)b.n) return -1;
  uint64_t ki = ((uint64_t)b.morton[i] << 32) | (uint64_t)b.order[i];
  uint64_t kj = ((uint64_t)b.morton[j] << 32) | (uint64_t)b.order[j];
  return __clzll((long long)(ki ^ kj));
}

__global__ void k_lbvh_hierarchy(LbvhBuild b) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int ni = (int)b.n - 1;
  if (i >= ni) return;
  int d = (lbvh_delta(b, i, i + 1) - lbvh_delta(b, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = lbvh_delta(b, i, i - d);
  int lmax = 2;
  while (lbvh_delta(b, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (lbvh_delta(b, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = lbvh_delta(b, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (lbvh_delta(b, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? d : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  uint32_t left = (lo == gamma) ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
  uint32_t right = (hi == gamma + 1) ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  b.child[2 * i] = left;
  b.child[2 * i + 1] = right;
  b.rangeLo[i] = (uint32_t)lo;
  b.rangeHi[i] = (uint32_t)hi;
  b.parent[(left & kLeafBit) ? (uint32_t)ni + (left & ~kLeafBit) : left] = (uint32_t)i;
  b.parent[(right & kLeafBit) ? (uint32_t)ni + (right & ~kLeafBit) : right] = (uint32_t)i;
  b.visits[i] = 0;
}

__device__ __forceinline__ void child_box(const LbvhBuild& b, uint32_t c, vec3& lo, vec3& hi) {
  if (c & kLeafBit) {
    const BvhTri& t = b.tris[c & ~kLeafBit];
    vec3 a{t.v0[0], t.v0[1], t.v0[2]}, e{t.v1[0], t.v1[1], t.v1[2]}, f{t.v2[0], t.v2[1], t.v2[2]};
    lo = vmin(a, vmin(e, f));
    hi = vmax(a, vmax(e, f));
  } else {
    const float* nb = b.nodeBox + 6 * c;
    lo = vec3{nb[0], nb[1], nb[2]};
    hi = vec3{nb[3], nb[4], nb[5]};
  }
}

// Leaf triangles in sorted order, then the bottom-up refit from every leaf.
__global__ void k_lbvh_leaves(LbvhBuild b) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  uint32_t t = b.order[k];
  vec3 a, c, d;
  tri_verts(b, t, a, c, d);
  BvhTri& o = b.tris[k];
  o.v0[0] = a.x; o.v0[1] = a.y; o.v0[2] = a.z; o.prim = t;
  o.v1[0] = c.x; o.v1[1] = c.y; o.v1[2] = c.z; o._p1 = 0.f;
  o.v2[0] = d.x; o.v2[1] = d.y; o.v2[2] = d.z; o._p2 = 0.f;
}

__global__ void k_lbvh_refit(LbvhBuild b) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b.n) return;
  uint32_t ni = b.n - 1;
  uint32_t node = b.parent[ni + k];
  while (true) {
    // Agent-scope release of this subtree's box, then acquire of the sibling's: the
    // fences make the boxes visible across the XCDs' separate L2 caches.
    __threadfence();
    uint32_t prev = atomicAdd(&b.visits[node], 1u);
    if (prev == 0) return;  // the sibling subtree is not finished yet
    __threadfence();
    uint32_t l = b.child[2 * node], r = b.child[2 * node + 1];
    vec3 llo, lhi, rlo, rhi;
    child_box(b, l, llo, lhi);
    child_box(b, r, rlo, rhi);
    BvhNode& n = b.nodes[node];
    n.lmin[0] = llo.x; n.lmin[1] = llo.y; n.lmin[2] = llo.z;
    n.lmax[0] = lhi.x; n.lmax[1] = lhi.y; n.lmax[2] = lhi.z;
    n.rmin[0] = rlo.x; n.rmin[1] = rlo.y; n.rmin[2] = rlo.z;
    n.rmax[0] = rhi.x; n.rmax[1] = rhi.y; n.rmax[2] = rhi.z;
    vec3 lo = vmin(llo, rlo), hi = vmax(lhi, rhi);
    float* nb = b.nodeBox + 6 * node;
    nb[0] = lo.x;
    nb[1] = lo.y;
    nb[2] = lo.z;
    nb[3] = hi.x;
    nb[4] = hi.y;
    nb[5] = hi.z;
    if (node == 0) return;
    node = b.parent[node];
  }
}

__device__ __forceinline__ uint32_t encode_child(const LbvhBuild& b, uint32_t c) {
  if (c & kLeafBit) return kLeafBit | ((c & ~kLeafBit) << 3);
  uint32_t cnt = b.rangeHi[c] - b.rangeLo[c] + 1;
  if (cnt <= kLeafMax) return kLeafBit | (b.rangeLo[c] << 3) | (cnt - 1);
  return c;
}

__global__ void k_lbvh_encode(LbvhBuild b) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n - 1) return;
  b.nodes[i].left = encode_child(b, b.child[2 * i]);
  b.nodes[i].right = encode_child(b, b.child[2 * i + 1]);
  b.nodes[i]._pad[0] = 0;
  b.nodes[i]._pad[1] = 0;
}

// ---------------------------------------------------------------- 4-wide collapse
//
// The binary tree becomes a 4-wide tree level by level from the root.  A 4-wide node
// over binary node r starts from r's two children and repeatedly opens the internal
// child of largest surface area (the greedy SAH collapse) until it has four children or
// only leaves.  Its internal children become the next level's work items; their slots
// come from an exclusive scan of the per-item counts, so the layout is deterministic
// and each level is contiguous (breadth-first order).

struct Bvh4Collapse {
  const BvhNode* bin;
  Bvh4Node* out;
  const uint32_t* workIn;  // binary node of each item of this level
  uint32_t* workOut;       // binary node of each item of the next level
  uint32_t count;          // items in this level
  uint32_t levelBase;      // 4-wide index of this level's first node
  uint32_t* cnt;           // count + 1 entries: internal children per item (last = 0)
  const uint32_t* offs;    // exclusive scan of cnt
  uint32_t* tmpRef;        // 4 per item
  float* tmpBox;           // 24 per item: lo.xyz, hi.xyz per child
  uint32_t* tmpSrc;        // 4 per item: source node of each child (binary _pad fields)
  uint32_t* src;           // 4 per 4-wide node: source node of each child slot, for refits
};

__device__ __forceinline__ float half_area(vec3 lo, vec3 hi) {
  vec3 e = hi - lo;
  return e.x * e.y + e.y * e.z + e.z * e.x;
}

__global__ void k_bvh4_select(Bvh4Collapse c) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == c.count) c.cnt[i] = 0;
  if (i >= c.count) return;
  uint32_t ref[4], src[4];
  vec3 lo[4], hi[4];
  const BvhNode& r = c.bin[c.workIn[i]];
  ref[0] = r.left;
  src[0] = r._pad[0];
  lo[0] = ld3(r.lmin);
  hi[0] = ld3(r.lmax);
  ref[1] = r.right;
  src[1] = r._pad[1];
  lo[1] = ld3(r.rmin);
  hi[1] = ld3(r.rmax);
  uint32_t n = 2;
  while (n < 4) {
    int best = -1;
    float bestA = -1.f;
    for (uint32_t k = 0; k < n; k++) {
      if (ref[k] & kLeafBit) continue;
      float a = half_area(lo[k], hi[k]);
      if (a > bestA) {
        bestA = a;
        best = (int)k;
      }
    }
    if (best < 0) break;
    const BvhNode& o = c.bin[ref[best]];
    ref[best] = o.left;
    src[best] = o._pad[0];
    lo[best] = ld3(o.lmin);
    hi[best] = ld3(o.lmax);
    ref[n] = o.right;
    src[n] = o._pad[1];
    lo[n] = ld3(o.rmin);
    hi[n] = ld3(o.rmax);
    n++;
  }
  uint32_t internal = 0;
  for (uint32_t k = 0; k < 4; k++) {
    bool used = k < n;
    c.tmpRef[4 * i + k] = used ? ref[k] : kLeafBit;
    c.tmpSrc[4 * i + k] = used ? src[k] : 0u;
    float* b = c.tmpBox + 24 * i + 6 * k;
    b[0] = used ? lo[k].x : 0.f;
    b[1] = used ? lo[k].y : 0.f;
    b[2] = used ? lo[k].z : 0.f;
    b[3] = used ? hi[k].x : 0.f;
    b[4] = used ? hi[k].y : 0.f;
    b[5] = used ? hi[k].z : 0.f;
    if (used && !(ref[k] & kLeafBit)) internal++;
  }
  c.cnt[i] = internal;
  c.out[c.levelBase + i].count = n;
}

__global__ void k_bvh4_emit(Bvh4Collapse c) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c.count) return;
  Bvh4Node& o = c.out[c.levelBase + i];
  uint32_t n = o.count;
  uint32_t next = c.levelBase + c.count + c.offs[i];  // first slot of this item's internal children
  uint32_t w = c.offs[i];
  for (uint32_t k = 0; k < 4; k++) {
    const float* b = c.tmpBox + 24 * i + 6 * k;
    o.lox[k] = b[0];
    o.loy[k] = b[1];
    o.loz[k] = b[2];
    o.hix[k] = b[3];
    o.hiy[k] = b[4];
    o.hiz[k] = b[5];
    uint32_t ref = c.tmpRef[4 * i + k];
    c.src[4 * (c.levelBase + i) + k] = c.tmpSrc[4 * i + k];
    if (k < n && !(ref & kLeafBit)) {
      c.workOut[w++] = ref;
      o.child[k] = next++;
    } else {
      o.child[k] = ref;
    }
  }
  o._pad[0] = o._pad[1] = o._pad[2] = 0;
}

}  // namespace rs
