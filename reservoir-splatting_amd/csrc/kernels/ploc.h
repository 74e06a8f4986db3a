// GPU BVH construction by parallel locally-ordered clustering (Meister and Bittner,
// "Parallel Locally-Ordered Clustering for Bounding Volume Hierarchy Construction",
// IEEE TVCG 24(3), 2018).
//
// The LBVH's spatial-median splits leave boxes that overlap more than a surface-area
// heuristic tree, so rays visit more nodes.  PLOC starts from the Morton-ordered triangles
// (one cluster each) and repeats, until one cluster is left:
//   1. every cluster finds the neighbour within kPlocRadius positions (in Morton order)
//      whose union with it has the smallest surface area;
//   2. mutual nearest neighbours merge into a new node, created by the lower position;
//   3. the cluster array is compacted in order.
// The pair order (area, lower position, higher position) is total, so the best pair of
// each iteration is mutual and every iteration merges at least once; merge slots and the
// compaction come from exclusive scans, so the tree does not depend on thread timing.
// Then a top-down pass (iterations in reverse creation order) gives every node its
// first position in depth-first leaf order, triangles are reordered so each subtree is a
// contiguous range, and subtrees of at most kLeafMax triangles become leaf ranges of the
// binary node array that the 4-wide collapse consumes.
#pragma once
#include "lbvh.h"

namespace rs {

// Search radius in Morton order.  On the 1080p field scene radii 2, 4, 6, 8, 12, 16 and 32
// traced the frame at 93.4, 94.3, 97.6, 97.2, 95.5, 96.2 and 96.2 fps: the effect is a
// tree-shape one of about 2% either way, and 8 is consistently ahead of 16 with a cheaper build.
#ifndef RS_PLOC_RADIUS
#define RS_PLOC_RADIUS 8
#endif
static constexpr int kPlocRadius = RS_PLOC_RADIUS;

struct PlocBuild {
  uint32_t n;          // triangles = leaves; internal node ids are n .. 2n-2
  float* box;          // 6 per node: lo.xyz, hi.xyz
  uint32_t* count;     // triangles under each node
  uint32_t* child;     // 2 per internal node, at 2 * (id - n)
  uint32_t* cl;        // node id of each cluster of this iteration (m entries)
  uint32_t* clOut;     // the next iteration's clusters
  uint32_t* nn;        // nearest-neighbour position of each cluster
  uint32_t* mergeFlag; // m + 1 entries (last 0): cluster creates a node
  uint32_t* mergeOffs; // exclusive scan of mergeFlag
  uint32_t* keepFlag;  // m + 1 entries (last 0): cluster survives compaction
  uint32_t* keepOffs;  // exclusive scan of keepFlag
  uint32_t* offset;    // first depth-first leaf position of each node
  uint32_t m;          // clusters in this iteration
  uint32_t nodeBase;   // id of the first node this iteration creates
  const BvhTri* trisIn;   // Morton order
  BvhTri* trisOut;        // depth-first order
  BvhNode* nodes;         // binary nodes, index id - n
};

__device__ __forceinline__ void ploc_load_box(const float* b, uint32_t k, vec3& lo, vec3& hi) {
  lo = vec3{b[6 * k], b[6 * k + 1], b[6 * k + 2]};
  hi = vec3{b[6 * k + 3], b[6 * k + 4], b[6 * k + 5]};
}
__device__ __forceinline__ void ploc_store_box(float* b, uint32_t k, vec3 lo, vec3 hi) {
  b[6 * k] = lo.x;
  b[6 * k + 1] = lo.y;
  b[6 * k + 2] = lo.z;
  b[6 * k + 3] = hi.x;
  b[6 * k + 4] = hi.y;
  b[6 * k + 5] = hi.z;
}

__global__ void k_ploc_init(PlocBuild p) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  const BvhTri& t = p.trisIn[k];
  vec3 a{t.v0[0], t.v0[1], t.v0[2]}, b{t.v1[0], t.v1[1], t.v1[2]}, c{t.v2[0], t.v2[1], t.v2[2]};
  ploc_store_box(p.box, k, vmin(a, vmin(b, c)), vmax(a, vmax(b, c)));
  p.count[k] = 1;
  p.cl[k] = k;
}

// Pair order: surface area of the union, then the lower and the higher position.
__device__ __forceinline__ bool ploc_pair_less(float a, uint32_t i, uint32_t j, float b, uint32_t k) {
  if (a != b) return a < b;
  uint32_t lo1 = i < j ? i : j, hi1 = i < j ? j : i;
  uint32_t lo2 = i < k ? i : k, hi2 = i < k ? k : i;
  return lo1 != lo2 ? lo1 < lo2 : hi1 < hi2;
}

__global__ void k_ploc_nearest(PlocBuild p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.m) return;
  vec3 lo, hi;
  ploc_load_box(p.box, p.cl[i], lo, hi);
  uint32_t jLo = i > (uint32_t)kPlocRadius ? i - (uint32_t)kPlocRadius : 0u;
  uint32_t jHi = i + (uint32_t)kPlocRadius < p.m - 1 ? i + (uint32_t)kPlocRadius : p.m - 1;
  float best = __builtin_inff();
  uint32_t bestJ = i;
  for (uint32_t j = jLo; j <= jHi; j++) {
    if (j == i) continue;
    vec3 l2, h2;
    ploc_load_box(p.box, p.cl[j], l2, h2);
    float a = half_area(vmin(lo, l2), vmax(hi, h2));
    if (bestJ == i || ploc_pair_less(a, i, j, best, bestJ)) {
      best = a;
      bestJ = j;
    }
  }
  p.nn[i] = bestJ;
}

__global__ void k_ploc_flags(PlocBuild p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == p.m) {
    p.mergeFlag[i] = 0;
    p.keepFlag[i] = 0;
  }
  if (i >= p.m) return;
  uint32_t j = p.nn[i];
  bool mutual = p.nn[j] == i;
  p.mergeFlag[i] = (mutual && i < j) ? 1u : 0u;
  p.keepFlag[i] = (mutual && i > j) ? 0u : 1u;
}

// Creates the merged nodes and writes the compacted cluster array (reads cl, writes clOut).
__global__ void k_ploc_merge(PlocBuild p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.m || !p.keepFlag[i]) return;
  uint32_t id = p.cl[i];
  if (p.mergeFlag[i]) {
    uint32_t L = id, R = p.cl[p.nn[i]];
    id = p.nodeBase + p.mergeOffs[i];
    p.child[2 * (id - p.n)] = L;
    p.child[2 * (id - p.n) + 1] = R;
    vec3 l1, h1, l2, h2;
    ploc_load_box(p.box, L, l1, h1);
    ploc_load_box(p.box, R, l2, h2);
    ploc_store_box(p.box, id, vmin(l1, l2), vmax(h1, h2));
    p.count[id] = p.count[L] + p.count[R];
  }
  p.clOut[p.keepOffs[i]] = id;
}

// Depth-first positions for the nodes created in one iteration (their parents are set).
__global__ void k_ploc_offsets(PlocBuild p, uint32_t first, uint32_t num) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num) return;
  uint32_t id = first + t;
  uint32_t L = p.child[2 * (id - p.n)], R = p.child[2 * (id - p.n) + 1];
  uint32_t off = p.offset[id];
  p.offset[L] = off;
  p.offset[R] = off + p.count[L];
}

__global__ void k_ploc_reorder(PlocBuild p) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  p.trisOut[p.offset[k]] = p.trisIn[k];
}

__device__ __forceinline__ uint32_t ploc_encode(const PlocBuild& p, uint32_t c) {
  uint32_t cnt = p.count[c];
  if (cnt <= kLeafMax) return kLeafBit | (p.offset[c] << 3) | (cnt - 1);
  return c - p.n;
}

// Binary nodes in the layout of the LBVH stage (child boxes in the parent).
__global__ void k_ploc_encode(PlocBuild p) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + 1 >= p.n) return;
  uint32_t L = p.child[2 * i], R = p.child[2 * i + 1];
  vec3 l1, h1, l2, h2;
  ploc_load_box(p.box, L, l1, h1);
  ploc_load_box(p.box, R, l2, h2);
  BvhNode& o = p.nodes[i];
  o.lmin[0] = l1.x; o.lmin[1] = l1.y; o.lmin[2] = l1.z;
  o.lmax[0] = h1.x; o.lmax[1] = h1.y; o.lmax[2] = h1.z;
  o.rmin[0] = l2.x; o.rmin[1] = l2.y; o.rmin[2] = l2.z;
  o.rmax[0] = h2.x; o.rmax[1] = h2.y; o.rmax[2] = h2.z;
  o.left = ploc_encode(p, L);
  o.right = ploc_encode(p, R);
  o._pad[0] = L;  // clustering node ids of the children, for refits
  o._pad[1] = R;
}

// ---------------------------------------------------------------- 4-wide collapse by SAH
//
// A 4-wide node costs the same to visit whatever its children are, so the expected cost of
// a collapse is proportional to the summed surface area of the binary nodes that become
// 4-wide nodes (every leaf range is a child slot of exactly one node, which adds the same
// cost to every collapse).  F(x, k) is the least cost of covering the subtree of clustering
// node x with at most k roots:
//   F(x, 1) = A(x) + D(x, 4),   F(x, k) = min(F(x, 1), D(x, k)) for k = 2..4,
//   D(x, k) = min over j = 1..k-1 of F(left, j) + F(right, k - j),
// and F = 0 for leaf ranges: the dynamic-programming conversion of Ylitalo, Karras and
// Laine (HPG 2017) with the leaf rule held fixed.  k_bvh4_cost fills F bottom-up, one
// clustering iteration per launch; k_bvh4_select_sah gives each 4-wide node its slots
// top-down by the minimising splits.  Ties take the smaller left share, so the tree is
// deterministic.

// With RS_BVH4_LEAF_COST > 0 the program also chooses the leaves: a subtree of at most
// kLeafCap triangles (contiguous in depth-first order) may become one leaf, at cost
// A(x) * count * RS_BVH4_LEAF_COST, a triangle test priced relative to a node visit, and
// F(x, 1) = min(that, A(x) + D(x, 4)).  The sign bit of the stored F(x, 1) records the
// leaf choice.  At 0 the fixed rule above holds (subtrees of at most kLeafMax triangles).
// On the 1080p field scene costs 0.15, 0.3, 0.5 and 0.8 traced the frame at 81.5, 92.6,
// 97.2 and 98.6 fps against 98.4 for the fixed rule, which stays the default.
#ifndef RS_BVH4_LEAF_COST
#define RS_BVH4_LEAF_COST 0
#endif
static constexpr float kLeafCost = RS_BVH4_LEAF_COST;
static constexpr uint32_t kLeafCap = 8;  // leaf codes hold counts 1..8

// Never opened: a triangle, or a leaf range of the fixed rule.
__device__ __forceinline__ bool dp_forced_leaf(const PlocBuild& p, uint32_t x) {
  return kLeafCost > 0.f ? x < p.n : p.count[x] <= kLeafMax;
}

// The 4-wide tree's reference for a child slot over clustering node x.
__device__ __forceinline__ uint32_t dp_ref(const PlocBuild& p, const float* F, uint32_t x) {
  const bool leaf = dp_forced_leaf(p, x) || (kLeafCost > 0.f && __builtin_signbit(F[4 * (size_t)(x - p.n)]));
  return leaf ? kLeafBit | (p.offset[x] << 3) | (p.count[x] - 1) : x - p.n;
}

__device__ __forceinline__ void dp_costs(const PlocBuild& p, const float* F, uint32_t x, float f[4]) {
  if (dp_forced_leaf(p, x)) {
    float c = 0.f;
    if (kLeafCost > 0.f) {
      vec3 lo, hi;
      ploc_load_box(p.box, x, lo, hi);
      c = half_area(lo, hi) * kLeafCost;
    }
    f[0] = f[1] = f[2] = f[3] = c;
    return;
  }
  const float* q = F + 4 * (size_t)(x - p.n);
  f[0] = fabsf(q[0]);
  f[1] = q[1];
  f[2] = q[2];
  f[3] = q[3];
}

// D(x, k) from the children's cost tables; `left` receives the minimising left share.
__device__ __forceinline__ float dp_split(const float fl[4], const float fr[4], uint32_t k, uint32_t& left) {
  float d = __builtin_inff();
  left = 1;
  for (uint32_t j = 1; j < k; j++) {
    float v = fl[j - 1] + fr[k - j - 1];
    if (v < d) {
      d = v;
      left = j;
    }
  }
  return d;
}

__global__ void k_bvh4_cost(PlocBuild p, float* F, uint32_t first, uint32_t num) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num) return;
  uint32_t id = first + t;
  if (dp_forced_leaf(p, id)) return;  // its table is never read
  float fl[4], fr[4];
  dp_costs(p, F, p.child[2 * (id - p.n)], fl);
  dp_costs(p, F, p.child[2 * (id - p.n) + 1], fr);
  vec3 lo, hi;
  ploc_load_box(p.box, id, lo, hi);
  uint32_t s;
  float d2 = dp_split(fl, fr, 2, s), d3 = dp_split(fl, fr, 3, s), d4 = dp_split(fl, fr, 4, s);
  const float area = half_area(lo, hi);
  float f1 = area + d4;
  bool leaf = false;
  if (kLeafCost > 0.f && p.count[id] <= kLeafCap) {
    const float c = area * (float)p.count[id] * kLeafCost;
    leaf = c <= f1;
    f1 = leaf ? c : f1;
  }
  float* q = F + 4 * (size_t)(id - p.n);
  q[0] = leaf ? -f1 : f1;
  q[1] = fminf(f1, d2);
  q[2] = fminf(f1, d3);
  q[3] = fminf(f1, d4);
}

__global__ void k_bvh4_select_sah(Bvh4Collapse c, PlocBuild p, const float* F) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == c.count) c.cnt[i] = 0;
  if (i >= c.count) return;
  // Pending (node, slot budget) pairs; the budgets of the pending pairs and the emitted
  // slots always sum to 4, so neither list outgrows four entries.
  uint32_t stNode[4], stBudget[4], slot[4];
  int sp = 0;
  uint32_t n = 0;
  {
    const uint32_t r = c.workIn[i] + p.n;  // the root of this node is always opened
    const uint32_t L = p.child[2 * (r - p.n)], R = p.child[2 * (r - p.n) + 1];
    float fl[4], fr[4];
    dp_costs(p, F, L, fl);
    dp_costs(p, F, R, fr);
    uint32_t k;
    dp_split(fl, fr, 4, k);
    stNode[0] = R;
    stBudget[0] = 4 - k;
    stNode[1] = L;
    stBudget[1] = k;
    sp = 2;
  }
  while (sp > 0) {
    sp--;
    const uint32_t x = stNode[sp], b = stBudget[sp];
    if (b > 1 && !dp_forced_leaf(p, x)) {
      const uint32_t L = p.child[2 * (x - p.n)], R = p.child[2 * (x - p.n) + 1];
      float fl[4], fr[4];
      dp_costs(p, F, L, fl);
      dp_costs(p, F, R, fr);
      uint32_t k;
      float d = dp_split(fl, fr, b, k);
      if (d < fabsf(F[4 * (size_t)(x - p.n)])) {
        stNode[sp] = R;
        stBudget[sp] = b - k;
        stNode[sp + 1] = L;
        stBudget[sp + 1] = k;
        sp += 2;
        continue;
      }
    }
    slot[n++] = x;
  }
  uint32_t internal = 0;
  for (uint32_t k = 0; k < 4; k++) {
    const bool used = k < n;
    const uint32_t x = used ? slot[k] : 0u;
    const uint32_t ref = used ? dp_ref(p, F, x) : kLeafBit;
    vec3 lo{0.f, 0.f, 0.f}, hi{0.f, 0.f, 0.f};
    if (used) ploc_load_box(p.box, x, lo, hi);
    c.tmpRef[4 * i + k] = ref;
    c.tmpSrc[4 * i + k] = x;
    float* bx = c.tmpBox + 24 * i + 6 * k;
    bx[0] = lo.x;
    bx[1] = lo.y;
    bx[2] = lo.z;
    bx[3] = hi.x;
    bx[4] = hi.y;
    bx[5] = hi.z;
    if (used && !(ref & kLeafBit)) internal++;
  }
  c.cnt[i] = internal;
  c.out[c.levelBase + i].count = n;
}

// ---------------------------------------------------------------- refit
//
// New vertex positions with the same topology: the tree is kept and every box is
// recomputed bottom-up (leaf triangles, clustering nodes in creation order, then the
// 4-wide nodes from the clustering node behind each child slot).  Hits stay exact; only
// the tree's quality drifts as geometry moves, until the next full build.

__global__ void k_refit_tris(BvhTri* tris, uint32_t n, const float* positions, const uint32_t* indices) {
  uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  BvhTri& t = tris[q];
  uint32_t prim = t.prim;
  vec3 a = load3(positions, indices[3 * prim]), b = load3(positions, indices[3 * prim + 1]),
       c = load3(positions, indices[3 * prim + 2]);
  t.v0[0] = a.x; t.v0[1] = a.y; t.v0[2] = a.z;
  t.v1[0] = b.x; t.v1[1] = b.y; t.v1[2] = b.z;
  t.v2[0] = c.x; t.v2[1] = c.y; t.v2[2] = c.z;
}

__global__ void k_refit_leaf_boxes(PlocBuild p) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.n) return;
  const BvhTri& t = p.trisOut[p.offset[k]];
  vec3 a{t.v0[0], t.v0[1], t.v0[2]}, b{t.v1[0], t.v1[1], t.v1[2]}, c{t.v2[0], t.v2[1], t.v2[2]};
  ploc_store_box(p.box, k, vmin(a, vmin(b, c)), vmax(a, vmax(b, c)));
}

__global__ void k_refit_nodes(PlocBuild p, uint32_t first, uint32_t num) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= num) return;
  uint32_t id = first + t;
  uint32_t L = p.child[2 * (id - p.n)], R = p.child[2 * (id - p.n) + 1];
  vec3 l1, h1, l2, h2;
  ploc_load_box(p.box, L, l1, h1);
  ploc_load_box(p.box, R, l2, h2);
  ploc_store_box(p.box, id, vmin(l1, l2), vmax(h1, h2));
}

__global__ void k_refit_bvh4(Bvh4Node* nodes, const uint32_t* src, const float* box, uint32_t count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  Bvh4Node& o = nodes[i];
  uint32_t n = o.count;
  for (uint32_t k = 0; k < n; k++) {
    vec3 lo, hi;
    ploc_load_box(box, src[4 * i + k], lo, hi);
    o.lox[k] = lo.x;
    o.loy[k] = lo.y;
    o.loz[k] = lo.z;
    o.hix[k] = hi.x;
    o.hiy[k] = hi.y;
    o.hiz[k] = hi.z;
  }
}

}  // namespace rs
