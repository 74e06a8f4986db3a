// MI355X ReservoirSplatting pass: kernels, LBVH build and the C ABI (include/rs.h).
//
// Host sequencing follows ReservoirSplatting::execute (RS/ReservoirSplatting.cpp:399-509):
// beginFrame (seed update :1629, keyframe count :1633-1641), camera keyframe ring
// (:1360-1388), initial candidates, the temporal mode switch (skipped on frame 0 and
// after reset, :423-465), spatial ping-pong (:1994-2020), resolve (:2022-2037) and
// endFrame's history swap (:1648-1680).  Option validation follows parseProperties /
// validateOptions (:192-301), frame-size checks setFrameDim (:356-375).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rs.h"
#include "kernels/ploc.h"
#include "pass/stages.h"
#include "pass/tasks.h"
#include "pass/adjoint.h"
#include "pass/accumulate.h"

using namespace rs;

static_assert(kQuirkRobustGrey == RS_QUIRK_ROBUST_GREY && kQuirkPathTimeZero == RS_QUIRK_PATH_TIME_ZERO &&
                  kQuirkMultiProjection == RS_QUIRK_MULTI_PROJECTION &&
                  kQuirkMisDropsLightHits == RS_QUIRK_MIS_DROPS_LIGHT_HITS &&
                  kQuirkSppLastSample == RS_QUIRK_SPP_LAST_SAMPLE,
              "device quirk bits follow include/rs.h");

// ------------------------------------------------------------------ kernels

namespace {

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
  // 64-lane butterfly reduction, then one atomic per wave.
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

struct KernelArgs {
  FrameParams p;
  SceneView s;
  BvhView bvh;
  PassBuffers b;
  TemporalOptions o;
  unsigned long long* rays;  // [0] closest, [1] visibility
};

// Pixel of this lane: a block covers a (1 << tileShift) x (256 >> tileShift) tile (16 x 16
// by default; the batch autotuner may pick another shape).  In the 16 x 16 tile each wave
// takes an 8 x 8 quadrant, lanes row-major inside it: a squarer patch than a 16 x 4 strip,
// whose rays share more of the tree (k_initial 4.56 -> 4.52 ms at 1080p on the field).
// Other shapes are row-major over the whole tile.
__device__ __forceinline__ bool tile_lane_pixel(const FrameParams& p, const TileId& t, uint32_t lane, uint32_t& x,
                                                uint32_t& y) {
  const uint32_t sh = p.tileShift;
  if (sh == 4) {
    const uint32_t w = lane >> 6, l = lane & 63u;
    x = (t.bx << 4) + ((w & 1u) << 3) + (l & 7u);
    y = (t.by << 4) + ((w >> 1) << 3) + (l >> 3);
    return x < p.W && y < p.H;
  }
  x = (t.bx << sh) + (lane & ((1u << sh) - 1u));
  y = t.by * (256u >> sh) + (lane >> sh);
  return x < p.W && y < p.H;
}
__device__ __forceinline__ bool lane_pixel(const FrameParams& p, const TileId& t, uint32_t& x, uint32_t& y) {
  return tile_lane_pixel(p, t, threadIdx.x, x, y);
}

// The kernel's first argument, read in place from the kernel-argument segment.  Binding a
// by-value struct argument to a reference instead lets the compiler copy it to per-lane
// scratch memory (k_initial: 1.4 KB more scratch per lane, 46 more spilled registers).
template <class T>
__device__ __forceinline__ const T& kernarg0() {
  return *(const T*)(const __attribute__((address_space(4))) T*)__builtin_amdgcn_kernarg_segment_ptr();
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Largest |coordinate| of n floats, as the bits of a non-negative float (whose order is the
// unsigned order of its bits): one atomic max per wave.
__global__ void k_max_abs_bits(const float* v, size_t n, uint32_t* out) {
  uint32_t m = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t b = __float_as_uint(v[i]) & 0x7fffffffu;
    if (b <= 0x7f800000u && b > m) m = b;  // NaN coordinates are left out
  }
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63u) == 0) atomicMax(out, m);
}

// Environment-map importance map: the base level, then each mip from the one above.
__global__ void k_env_importance(EnvMapView e, float* imp) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= kEnvImportanceDim) return;
  imp[(size_t)y * kEnvImportanceDim + x] = env_importance_texel(e, x, y);
}

__global__ void k_env_mip(const float* prev, float* out, uint32_t prevDim) {
  const uint32_t dim = prevDim / 2;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim * dim) return;
  out[i] = env_importance_mip_texel(prev, prevDim, i % dim, i / dim);
}

// Waves per SIMD the path-tracing and resampling kernels are compiled for.  Left to
// itself the compiler gives these kernels 170-250 VGPRs (2 waves per SIMD), too few to
// hide the latency of the BVH walk; capping the registers keeps rarely used path state
// in (cached) scratch instead.  Measured at 1080p on the 1M-triangle field: 2 -> 6 waves
// cut k_initial from 13.6 to 8.0 ms at the time, and 7 waves now takes it from 5.1 to
// 8.2 ms (spills); the temporal kernels run best at 5 (2.99 -> 2.94 ms against 4; the
// MultiScatter frame 53.3 -> 53.9 fps), the spatial kernel at 5 (4 and 6 are no faster).
#ifndef RS_WAVES_INITIAL
#define RS_WAVES_INITIAL 6
#endif
#ifndef RS_WAVES_TEMPORAL
#define RS_WAVES_TEMPORAL 5
#endif
#ifndef RS_WAVES_SPATIAL
#define RS_WAVES_SPATIAL 5
#endif
#define RS_HEAVY_KERNEL(w) __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(w)))

// The heavy kernels are instantiated twice: kEnv = false for scenes without an
// environment map, whose code then carries none of its branches (measured 4% faster
// k_initial at 1080p on the field), and kEnv = true.
//
// Every per-pixel kernel is a body over (arguments, tile) with two entry points: k_* takes
// its arguments by value and its tile from the block index; kb_* (batched frames, below)
// takes both from the batch table.
#define RS_TRACE_PROLOGUE_ENV(env)                        \
  __shared__ uint32_t s_stack[kStackLds * kBlock];        \
  __shared__ float s_dist[kStackLds * kBlock];            \
  Tracer tr;                                              \
  tr.bvh = a.bvh;                                         \
  tr.lds = (LdsU32*)(s_stack + threadIdx.x);              \
  tr.ldsDist = (LdsF32*)(s_dist + threadIdx.x);           \
  tr.rays_closest = 0;                                    \
  tr.rays_visibility = 0;                                 \
  DevCtx c{&a.p, &a.s, &tr, (env)};                       \
  uint32_t x, y;                                          \
  bool inside = lane_pixel(a.p, tile, x, y);
#define RS_TRACE_PROLOGUE RS_TRACE_PROLOGUE_ENV(a.s.env.present != 0u)
#define RS_TRACE_PROLOGUE_T RS_TRACE_PROLOGUE_ENV(kEnv && a.s.env.present != 0u)

#define RS_TRACE_EPILOGUE                                 \
  wave_add_u64(&a.rays[0], tr.rays_closest);              \
  wave_add_u64(&a.rays[1], tr.rays_visibility);

__device__ __forceinline__ void vbuffer_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE
  if (inside) primary_vbuffer(c, a.b, x, y);
}
__global__ __launch_bounds__(256) void k_vbuffer(KernelArgs) { vbuffer_body(kernarg0<KernelArgs>(), grid_tile()); }

// Initial candidates, one lane per pixel.  Lanes whose path ends early idle until the
// wave's longest path finishes; refilling them with further pixels (a persistent kernel
// that starts a new path in a lane as soon as its path ends) was measured slower at 1080p
// on the field (6.8 against 5.0 ms): the wave then mixes primary and deep segments, whose
// traversals and hit handling diverge and share less cached data.
// kStats: with a per-pixel ray-count output (PixelStats); the default instantiation carries
// none of its code, which would otherwise add register spills to the whole path loop.
template <bool kEnv, bool kStats>
__device__ __forceinline__ void initial_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE_T
  if (inside) {
    const FrameParams& p = a.p;
    for (uint32_t s = 0; s < p.spp; s++) {
      PathState ps;
      candidate_start(c, ps, candidate_path_id(p, x, y, s));
      bool primary = true;
      do {
        candidate_step(c, ps, primary);
        primary = false;
      } while (ps.active());
      candidate_finish(c, a.b, ps);
      // PixelStats: this kernel's rays belong to the pixel it traces (RS/PathTracer.slang:304-308);
      // the path's id names the pixel, so no pixel coordinate stays live across the loop.
      if (kStats && s + 1 == p.spp)
        a.b.rayCountOut[ps.pixelY() * p.W + ps.pixelX()] = tr.rays_closest + tr.rays_visibility;
    }
  }
  if (kStats) wave_add_u64(&a.rays[2], tr.rays_closest + tr.rays_visibility);
  RS_TRACE_EPILOGUE
}
template <bool kEnv, bool kStats>
RS_HEAVY_KERNEL(RS_WAVES_INITIAL) void k_initial(KernelArgs) {
  initial_body<kEnv, kStats>(kernarg0<KernelArgs>(), grid_tile());
}

// Next chunk of 64 tasks for this wave (the same value in every lane).
__device__ __forceinline__ uint32_t fetch_chunk(uint32_t* fetch) {
  uint32_t base = 0;
  if ((threadIdx.x & 63u) == 0) base = atomicAdd(fetch, 64u);
  return __shfl(base, 0, 64);
}

__device__ __forceinline__ void collect_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE
  if (inside) collect_temporal(c, a.b, a.o, x, y);
  RS_TRACE_EPILOGUE
}
__global__ __launch_bounds__(256) void k_collect(KernelArgs) { collect_body(kernarg0<KernelArgs>(), grid_tile()); }

__device__ __forceinline__ void reproject_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE
  if (inside) {
    uint32_t dst = reproject(c, a.b, x, y);
    a.b.splatDst[y * a.p.W + x] = dst;
    if (dst != kNoSplat) a.b.splatRank[y * a.p.W + x] = atomicAdd(&a.b.cellCount[dst], 1u);
  }
  RS_TRACE_EPILOGUE
}
__global__ __launch_bounds__(256) void k_reproject(KernelArgs) { reproject_body(kernarg0<KernelArgs>(), grid_tile()); }

// One lane per pixel over every partition: the batched form's MultiReproject (frames alone
// run the compacted k_multi_flags / k_multi_reproject_tasks).
template <bool kEnv>
__device__ __forceinline__ void multi_reproject_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE_T
  if (inside) multi_reproject(c, a.b, a.o.numPartitions, x, y);
  RS_TRACE_EPILOGUE
}

// Scatter source pixels into their cell segments (arbitrary order inside a segment: the
// ranks come from the reprojection's atomics, and k_splat_order fixes the order).
// Entry i is partition i / P, source pixel i % P.
// Entry i goes to its cell's segment at the rank its reprojection drew from the cell's count.
// sortedCell (the cell of each entry, for the candidate-compacted resampling) may be null.
__device__ __forceinline__ void splat_fill_entry(const PassBuffers& b, uint32_t* sortedCell, uint32_t i, uint32_t P) {
  uint32_t d = b.splatDst[i];
  if (d == kNoSplat) return;
  uint32_t slot = b.splatRank[i];
  b.sorted[b.cellOffset[d] + slot] = i % P;
  if (sortedCell) sortedCell[b.cellOffset[d] + slot] = d;
}
__global__ void k_splat_fill(PassBuffers b, uint32_t* sortedCell, uint32_t n, uint32_t P) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) splat_fill_entry(b, sortedCell, i, P);
}

// Order each cell segment by ascending source pixel.  A segment of at most kSortSmall
// entries (nearly all of them: a pixel receives 0.49 splats on average at 1080p) is
// insertion-sorted by its cell's thread; a longer one is listed for k_splat_order_large.
static constexpr uint32_t kSortSmall = 16;
static constexpr uint32_t kSortChunk = 2048;  // entries one block sorts in LDS at a time
static constexpr uint32_t kSortLargeBlocks = 256;  // blocks that share the long segments

__device__ __forceinline__ void splat_order_cell(const PassBuffers& b, uint32_t cell, uint32_t* largeCells,
                                                 uint32_t* largeCount) {
  uint32_t cnt = b.cellCount[cell];
  if (cnt < 2) return;
  if (cnt > kSortSmall) {
    largeCells[atomicAdd(largeCount, 1u)] = cell;
    return;
  }
  uint32_t* seg = b.sorted + b.cellOffset[cell];
  for (uint32_t i = 1; i < cnt; i++) {
    uint32_t v = seg[i];
    uint32_t j = i;
    while (j > 0 && seg[j - 1] > v) {
      seg[j] = seg[j - 1];
      j--;
    }
    seg[j] = v;
  }
}
__global__ void k_splat_order(PassBuffers b, uint32_t n, uint32_t* largeCells, uint32_t* largeCount) {
  uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell < n) splat_order_cell(b, cell, largeCells, largeCount);
}

// Number of entries of the ascending run r[0, len) below v.
__device__ __forceinline__ uint32_t rank_below(const uint32_t* r, uint32_t len, uint32_t v) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (r[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The long segments, one block per listed cell: each chunk of kSortChunk entries is sorted
// in LDS by a bitonic network, then runs of width kSortChunk, 2 kSortChunk, ... are merged
// pairwise, every entry placed at (its rank in its own run) + (its rank in the partner run),
// ping-ponging between the segment and the same range of `temp` (the splat ranks, free once
// the fill is done; they have an entry per splat-list position).  The entries of a cell are distinct source pixels, so the
// order is unique: O(k log^2 k) work for a segment of k entries, whatever k.
__device__ __forceinline__ void splat_order_large_cell(const PassBuffers& b, uint32_t cell, uint32_t* temp, uint32_t* s) {
  const uint32_t cnt = b.cellCount[cell], off = b.cellOffset[cell];
  uint32_t* seg = b.sorted + off;
  for (uint32_t c0 = 0; c0 < cnt; c0 += kSortChunk) {
    const uint32_t len = min(kSortChunk, cnt - c0);
    for (uint32_t i = threadIdx.x; i < kSortChunk; i += blockDim.x) s[i] = i < len ? seg[c0 + i] : 0xffffffffu;
    __syncthreads();
    for (uint32_t k = 2; k <= kSortChunk; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < kSortChunk; i += blockDim.x) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const uint32_t x = s[i], y = s[ixj];
            const bool up = (i & k) == 0;
            if ((x > y) == up) {
              s[i] = y;
              s[ixj] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) seg[c0 + i] = s[i];
    __syncthreads();
  }
  uint32_t* src = seg;
  uint32_t* dst = temp + off;
  for (uint32_t w = kSortChunk; w < cnt; w <<= 1) {
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
      const uint32_t run = i / w, pair = (run & ~1u) * w;
      const uint32_t v = src[i];
      uint32_t pos;
      if ((run & 1u) == 0) {  // left run: the partner is the run after it (possibly empty)
        const uint32_t bStart = min(pair + w, cnt), bLen = min(pair + 2 * w, cnt) - bStart;
        pos = (i - pair) + rank_below(src + bStart, bLen, v);
      } else {                // right run: the partner is the full run before it
        pos = (i - pair - w) + rank_below(src + pair, w, v);
      }
      dst[pair + pos] = v;
    }
    __syncthreads();
    uint32_t* t = src;
    src = dst;
    dst = t;
  }
  if (src != seg) {
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) seg[i] = src[i];
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_splat_order_large(PassBuffers b, const uint32_t* largeCells,
                                                           const uint32_t* largeCount, uint32_t* temp) {
  __shared__ uint32_t s[kSortChunk];
  const uint32_t nLarge = *largeCount;
  for (uint32_t li = blockIdx.x; li < nLarge; li += gridDim.x) splat_order_large_cell(b, largeCells[li], temp, s);
}

// ---- candidate-compacted temporal resampling (pass/tasks.h)

enum TemporalTaskKind : uint32_t { kTasksScatter = 0, kTasksMulti = 1, kTasksBackup = 2 };

// Live flags of each pixel's current (and backup) candidate.
__device__ __forceinline__ void temporal_flags_body(const KernelArgs& a, const TileId& tile, const TaskLists& t,
                                                    bool withBackup) {
  uint32_t x, y;
  const bool inside = lane_pixel(a.p, tile, x, y);
  const uint32_t idx = inside ? y * a.p.W + x : 0u, pos = tile_position(tile);
  queue_flag(t.q, 0, pos, inside && any_pos(ldR(a.b.currR, idx).integrand));
  if (withBackup) queue_flag(t.q, 1, pos, inside && any_pos(ldR(a.b.backupR, idx).integrand));
}
__global__ __launch_bounds__(256) void k_temporal_flags(KernelArgs, TaskLists t, uint32_t withBackup) {
  temporal_flags_body(kernarg0<KernelArgs>(), grid_tile(), t, withBackup != 0);
}

// RobustReuseOptimization over the compacted live previous samples: the flags list every
// pixel whose previous sample has a positive integrand component; each task of the
// persistent kernel shifts one listed sample to one pixel of its 3x3 ring (8 consecutive
// lanes share a sample, so their walks start alike).
__device__ __forceinline__ void robust_flags_body(const KernelArgs& a, const TileId& tile, const TaskQueue& q) {
  uint32_t x, y;
  const bool inside = lane_pixel(a.p, tile, x, y);
  const uint32_t idx = inside ? y * a.p.W + x : 0u;
  queue_flag(q, 0, tile_position(tile), inside && any_pos(ldR(a.b.prevR, idx).integrand));
}
__global__ __launch_bounds__(256) void k_robust_flags(KernelArgs, TaskQueue q) {
  robust_flags_body(kernarg0<KernelArgs>(), grid_tile(), q);
}

// The persistent task kernels: a tracer without a pixel of its own.
#define RS_TASK_PROLOGUE                                  \
  __shared__ uint32_t s_stack[kStackLds * kBlock];        \
  __shared__ float s_dist[kStackLds * kBlock];            \
  Tracer tr;                                              \
  tr.lds = (LdsU32*)(s_stack + threadIdx.x);              \
  tr.ldsDist = (LdsF32*)(s_dist + threadIdx.x);           \
  tr.rays_closest = 0;                                    \
  tr.rays_visibility = 0;

// Takes chunks of 64 tasks until the list is exhausted, then adds the wave's rays.
__device__ __forceinline__ void robust_tasks_run(const DevCtx& c, const KernelArgs& a, const TaskQueue& q) {
  const uint32_t total = 8u * queue_start(q, 1);
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t base = fetch_chunk(q.fetch);
    if (base >= total) break;
    const uint32_t i = base + lane;
    if (i < total) robust_shift_task(c, a.b, q.queue[i >> 3], i & 7u);
  }
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void k_robust_tasks(KernelArgs, TaskQueue q) {
  const KernelArgs& a = kernarg0<KernelArgs>();
  RS_TASK_PROLOGUE
  tr.bvh = a.bvh;
  DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
  robust_tasks_run(c, a, q);
  RS_TRACE_EPILOGUE
}

// MultiReproject over the compacted live previous samples: the flags clear every pixel's K
// destinations and list the pixels whose previous sample is non-zero (the only ones that
// splat); each task of the persistent kernel reprojects one listed sample at one time
// partition, K consecutive lanes per sample, so that their walks start alike.
__global__ __launch_bounds__(256) void k_multi_flags(KernelArgs, TaskQueue q, uint32_t K) {
  const KernelArgs& a = kernarg0<KernelArgs>();
  const TileId tile = grid_tile();
  uint32_t x, y;
  const bool inside = lane_pixel(a.p, tile, x, y);
  const uint32_t P = a.p.W * a.p.H, idx = inside ? y * a.p.W + x : 0u;
  bool live = false;
  if (inside) {
    for (uint32_t i = 0; i < K; i++) a.b.splatDst[i * P + idx] = kNoSplat;
    live = !all_zero(ldR(a.b.prevR, idx).integrand);
  }
  queue_flag(q, 0, tile_position(tile), live);
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void k_multi_reproject_tasks(KernelArgs, TaskQueue q, uint32_t K) {
  const KernelArgs& a = kernarg0<KernelArgs>();
  RS_TASK_PROLOGUE
  tr.bvh = a.bvh;
  DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
  const uint32_t total = K * queue_start(q, 1);
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t base = fetch_chunk(q.fetch);
    if (base >= total) break;
    const uint32_t i = base + lane;
    if (i < total) multi_reproject_task(c, a.b, K, q.queue[i / K], i % K);
  }
  RS_TRACE_EPILOGUE
}

// Persistent: each wave takes 64 consecutive tasks at a time (current tasks, then backup
// tasks, then splat-list entries) until the lists are exhausted.  `part` selects the tasks:
// every task (kTasksAll), the current and backup lists (kTasksLists), or the splat-list
// entries alone (kTasksSplats, counted by t.splatFetch), which read nothing of the current
// frame's initial candidates and so can run while those are traced.
enum TemporalTaskPart : uint32_t { kTasksAll = 0, kTasksLists = 1, kTasksSplats = 2 };

__device__ __forceinline__ void temporal_tasks_run(const DevCtx& c, const KernelArgs& a, const TaskLists& t,
                                                   uint32_t kind, uint32_t part = kTasksAll) {
  const uint32_t P = a.p.W * a.p.H;
  const uint32_t splats = part == kTasksLists ? 0u : splat_total(a.b, P * a.o.numPartitions);
  const uint32_t nc = part == kTasksSplats ? 0u : queue_start(t.q, 1);
  const uint32_t nq = part == kTasksSplats ? 0u : queue_start(t.q, t.q.lists);
  const uint32_t total = nq + splats;
  uint32_t* fetch = part == kTasksSplats ? t.splatFetch : t.q.fetch;
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t base = fetch_chunk(fetch);
    if (base >= total) break;
    const uint32_t i = base + lane;
    if (i < nc) {
      if (kind == kTasksBackup) backup_current_task(c, a.b, a.o, t, t.q.queue[i]);
      else scatter_current_task(c, a.b, a.o, t, kind == kTasksMulti, t.q.queue[i]);
    } else if (i < nq) {
      backup_backup_task(c, a.b, a.o, t, t.q.queue[i]);
    } else if (i < total) {
      if (kind == kTasksBackup) backup_splat_task(c, a.b, a.o, t, i - nq);
      else scatter_splat_task(c, a.b, a.o, t, kind == kTasksMulti, i - nq);
    }
  }
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void k_temporal_tasks(KernelArgs, TaskLists t, uint32_t kind, uint32_t part) {
  const KernelArgs& a = kernarg0<KernelArgs>();
  RS_TASK_PROLOGUE
  tr.bvh = a.bvh;
  DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
  temporal_tasks_run(c, a, t, kind, part);
  RS_TRACE_EPILOGUE
}

__device__ __forceinline__ void temporal_combine_body(const KernelArgs& a, const TileId& tile, const TaskLists& t,
                                                      uint32_t kind) {
  uint32_t x, y;
  if (!lane_pixel(a.p, tile, x, y)) return;
  if (kind == kTasksBackup) backup_combine(a.p, a.b, a.o, t, x, y);
  else scatter_combine(a.p, a.b, a.o, t, kind == kTasksMulti, x, y);
}
__global__ __launch_bounds__(256) void k_temporal_combine(KernelArgs, TaskLists t, uint32_t kind) {
  temporal_combine_body(kernarg0<KernelArgs>(), grid_tile(), t, kind);
}

template <bool kEnv>
__device__ __forceinline__ void gather_body(const KernelArgs& a, const TileId& tile) {
  RS_TRACE_PROLOGUE_T
  if (inside) gather_temporal(c, a.b, a.o, x, y);
  RS_TRACE_EPILOGUE
}
template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void k_gather(KernelArgs) {
  gather_body<kEnv>(kernarg0<KernelArgs>(), grid_tile());
}

// ---- candidate-compacted spatial reuse (pass/tasks.h)

__device__ __forceinline__ void spatial_flags_body(const KernelArgs& a, const TileId& tile, const SpatialTasks& t) {
  uint32_t x, y;
  const bool inside = lane_pixel(a.p, tile, x, y);
  spatial_flags(a.p, a.b, a.o, t, inside, x, y, tile_position(tile));
}
__global__ __launch_bounds__(256) void k_spatial_flags(KernelArgs, SpatialTasks t) {
  spatial_flags_body(kernarg0<KernelArgs>(), grid_tile(), t);
}

// Persistent like k_temporal_tasks over lists 0 .. 2n-1: list i the central sample into
// neighbour i, list n + i neighbour i's sample into the pixel.
__device__ __forceinline__ void spatial_tasks_run(const DevCtx& c, const KernelArgs& a, const SpatialTasks& t) {
  const uint32_t n = a.o.spatialNeighbors;
  const uint32_t total = queue_start(t.q, t.q.lists);
  const uint32_t lane = threadIdx.x & 63u;
  for (;;) {
    const uint32_t base = fetch_chunk(t.q.fetch);
    if (base >= total) break;
    const uint32_t i = base + lane;
    if (i < total) {
      uint32_t l = 0;
      while (i >= queue_start(t.q, l + 1)) l++;
      spatial_task(c, a.b, a.o, t, l / n, l % n, t.q.queue[i]);
    }
  }
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_SPATIAL) void k_spatial_tasks(KernelArgs, SpatialTasks t) {
  const KernelArgs& a = kernarg0<KernelArgs>();
  RS_TASK_PROLOGUE
  tr.bvh = a.bvh;
  DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
  spatial_tasks_run(c, a, t);
  RS_TRACE_EPILOGUE
}

__device__ __forceinline__ void spatial_combine_body(const KernelArgs& a, const TileId& tile, const SpatialTasks& t) {
  uint32_t x, y;
  if (lane_pixel(a.p, tile, x, y)) spatial_combine(a.p, a.b, a.o, t, x, y);
}
__global__ __launch_bounds__(256) void k_spatial_combine(KernelArgs, SpatialTasks t) {
  spatial_combine_body(kernarg0<KernelArgs>(), grid_tile(), t);
}

// ------------------------------------------------------------------ frozen replay

struct ReplayArgs {
  FrameParams p;
  SceneView s;
  const PackedRecon* C;       // frozen reconnection records (the frame's final reservoirs)
  FrozenPixel* frozen;
  const float* params;        // 4 per material
  uint32_t numMaterials;
};

__global__ __launch_bounds__(256) void k_freeze(ReplayArgs a, const float* color) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.p.W * a.p.H) return;
  ReconnectionData rd = ldC(a.C, i);
  ReplayVertex v1, v2;
  replay_vertices(a.p, a.s, rd, v1, v2);
  FrozenPixel fp;
  fp.color[0] = color[4 * i];
  fp.color[1] = color[4 * i + 1];
  fp.color[2] = color[4 * i + 2];
  fp.flags = (v1.active ? 1u : 0u) | (v2.active ? 2u : 0u);
  BSDFGrad g1{v3s(0.f), v3s(0.f), v3s(0.f)}, g2 = g1;
  if (v1.active) g1 = replay_eval(a.s, v1, a.params);
  if (v2.active) g2 = replay_eval(a.s, v2, a.params);
  fp.f1[0] = g1.f.x; fp.f1[1] = g1.f.y; fp.f1[2] = g1.f.z;
  fp.f2[0] = g2.f.x; fp.f2[1] = g2.f.y; fp.f2[2] = g2.f.z;
  fp.mat1 = v1.active ? v1.sd.materialID : 0u;
  fp.mat2 = v2.active ? v2.sd.materialID : 0u;
  a.frozen[i] = fp;
}

// R(theta) of every pixel (RGBA, alpha 1).
__global__ __launch_bounds__(256) void k_replay_forward(ReplayArgs a, float* image) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.p.W * a.p.H) return;
  FrozenPixel fp = a.frozen[i];
  float r[3] = {fp.color[0], fp.color[1], fp.color[2]};
  if (fp.flags) {
    ReconnectionData rd = ldC(a.C, i);
    ReplayVertex v1, v2;
    replay_vertices(a.p, a.s, rd, v1, v2);
    if (fp.flags & 1u) {
      vec3 f = replay_eval(a.s, v1, a.params).f;
      r[0] *= replay_ratio(f.x, fp.f1[0]);
      r[1] *= replay_ratio(f.y, fp.f1[1]);
      r[2] *= replay_ratio(f.z, fp.f1[2]);
    }
    if (fp.flags & 2u) {
      vec3 f = replay_eval(a.s, v2, a.params).f;
      r[0] *= replay_ratio(f.x, fp.f2[0]);
      r[1] *= replay_ratio(f.y, fp.f2[1]);
      r[2] *= replay_ratio(f.z, fp.f2[2]);
    }
  }
  image[4 * i] = r[0];
  image[4 * i + 1] = r[1];
  image[4 * i + 2] = r[2];
  image[4 * i + 3] = 1.f;
}

__global__ void k_zero_f64(double* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static constexpr uint32_t kReplayLdsMaterials = 256;

// Wave-level reduction of per-lane gradient contributions to one material each: the lanes
// sharing the first remaining lane's material sum their four values with a butterfly, and
// one lane adds the sums to `grad`; repeated until every active lane is done (once per
// distinct material in the wave, typically a few for an 8x8 pixel patch).
__device__ __forceinline__ void wave_add_by_material(uint32_t mat, bool active, const double v[4], double* grad) {
  const int lane = (int)(threadIdx.x & 63);
  uint64_t remaining = __ballot(active);
  while (remaining) {
    const int leader = __ffsll((unsigned long long)remaining) - 1;
    const uint32_t m = __shfl(mat, leader, 64);
    const bool mine = active && mat == m;
    double s[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      double x = mine ? v[k] : 0.0;
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
      s[k] = x;
    }
    if (lane == leader) {
#pragma unroll
      for (int k = 0; k < 4; k++)
        if (s[k] != 0.0) atomicAdd(&grad[4 * m + k], s[k]);
    }
    remaining &= ~__ballot(mine);
    if (mine) active = false;
  }
}

// dL/dtheta = sum over pixels and channels of dL/dR_c * dR_c/dtheta, in double.  With at
// most kReplayLdsMaterials materials a block first sums its pixels in LDS; otherwise each
// wave sums its lanes per material before the global atomics.
__global__ __launch_bounds__(256) void k_replay_backward(ReplayArgs a, const float* gradImage, double* grad) {
  __shared__ double s_grad[4 * kReplayLdsMaterials];
  const bool useLds = a.numMaterials <= kReplayLdsMaterials;
  if (useLds) {
    for (uint32_t k = threadIdx.x; k < 4 * a.numMaterials; k += blockDim.x) s_grad[k] = 0.0;
    __syncthreads();
  }
  // this lane's contributions to the materials at x1 and x2
  double v1[4] = {0.0, 0.0, 0.0, 0.0}, v2[4] = {0.0, 0.0, 0.0, 0.0};
  uint32_t m1 = 0, m2 = 0;
  bool act1 = false, act2 = false;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.p.W * a.p.H) {
    FrozenPixel fp = a.frozen[i];
    if (fp.flags) {
      ReconnectionData rd = ldC(a.C, i);
      ReplayVertex rv1, rv2;
      replay_vertices(a.p, a.s, rd, rv1, rv2);
      BSDFGrad g1{v3s(0.f), v3s(0.f), v3s(0.f)}, g2 = g1;
      if (fp.flags & 1u) g1 = replay_eval(a.s, rv1, a.params);
      if (fp.flags & 2u) g2 = replay_eval(a.s, rv2, a.params);
      const float f1v[3] = {g1.f.x, g1.f.y, g1.f.z}, f2v[3] = {g2.f.x, g2.f.y, g2.f.z};
      const float d1b[3] = {g1.dBase.x, g1.dBase.y, g1.dBase.z}, d1r[3] = {g1.dRough.x, g1.dRough.y, g1.dRough.z};
      const float d2b[3] = {g2.dBase.x, g2.dBase.y, g2.dBase.z}, d2r[3] = {g2.dRough.x, g2.dRough.y, g2.dRough.z};
      m1 = fp.mat1;
      m2 = fp.mat2;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float up = gradImage[4 * i + c] * fp.color[c];
        if (up == 0.f) continue;
        float r1 = (fp.flags & 1u) ? replay_ratio(f1v[c], fp.f1[c]) : 1.f;
        float r2 = (fp.flags & 2u) ? replay_ratio(f2v[c], fp.f2[c]) : 1.f;
        if ((fp.flags & 1u) && fp.f1[c] != 0.f) {
          float w = up * r2 / fp.f1[c];
          v1[c] += (double)(w * d1b[c]);
          v1[3] += (double)(w * d1r[c]);
          act1 = true;
        }
        if ((fp.flags & 2u) && fp.f2[c] != 0.f) {
          float w = up * r1 / fp.f2[c];
          v2[c] += (double)(w * d2b[c]);
          v2[3] += (double)(w * d2r[c]);
          act2 = true;
        }
      }
    }
  }
  if (useLds) {
    for (int k = 0; k < 4; k++) {
      if (act1 && v1[k] != 0.0) atomicAdd(&s_grad[4 * m1 + k], v1[k]);
      if (act2 && v2[k] != 0.0) atomicAdd(&s_grad[4 * m2 + k], v2[k]);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 4 * a.numMaterials; k += blockDim.x)
      if (s_grad[k] != 0.0) atomicAdd(&grad[k], s_grad[k]);
  } else {
    wave_add_by_material(m1, act1, v1, grad);
    wave_add_by_material(m2, act2, v2, grad);
  }
}

// VisualizeForwardReprojection::execute: splats received per pixel.
__global__ void k_scatter_counts(const uint32_t* cellCount, uint32_t P, uint32_t K, float* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  float cnt = 0.f;
  if (K > 1) {
    float w = 1.f / (float)K;
    for (uint32_t k = 0; k < K; k++) cnt += w * (float)cellCount[k * P + i];
  } else {
    cnt = (float)cellCount[i];
  }
  out[i] = cnt;
}

__global__ __launch_bounds__(256) void k_resolve(FrameParams p, PassBuffers b, const PackedReservoir* R) {
  uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x < p.W && y < p.H) resolve(p, b, R, x, y);
}

// Instance motion: world vertices of every instanced vertex from its object-space vertex.
__global__ void k_instance_transform(const float* objP, const float* objN, const uint32_t* vertInstance,
                                     const float* xforms, uint32_t nv, float* P, float* N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const uint32_t k = vertInstance[i];
  if (k == kNoInstance) return;
  const float* m = xforms + 12 * (size_t)k;
  const vec3 p = instance_point(m, load3(objP, i)), n = instance_normal(m, load3(objN, i));
  P[3 * i] = p.x;
  P[3 * i + 1] = p.y;
  P[3 * i + 2] = p.z;
  N[3 * i] = n.x;
  N[3 * i + 1] = n.y;
  N[3 * i + 2] = n.z;
}

// The light collection's triangles at the current positions, with the arithmetic of the
// host build (build_light_collection); flux[i] < 0 marks a triangle that became degenerate.
__global__ void k_update_lights(EmissiveTriangle* lights, uint32_t n, const float* P, const uint32_t* indices,
                                float* flux) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  EmissiveTriangle e = lights[i];
  const uint32_t t = e.prim;
  e.p0 = load3(P, indices[3 * t]);
  e.p1 = load3(P, indices[3 * t + 1]);
  e.p2 = load3(P, indices[3 * t + 2]);
  const vec3 cr = cross(e.p1 - e.p0, e.p2 - e.p0);
  const float len = length(cr);
  if (!(len > 0.f)) {
    flux[i] = -1.f;
    return;
  }
  e.area = 0.5f * len;
  e.normal = cr / len;
  e.flux = luminance(e.Le) * e.area * kPi;
  lights[i] = e;
  flux[i] = e.flux;
}

// Per-pixel ray counts: the initial-candidate kernel wrote each pixel's own rays; every
// other kernel's rays belong to pixel (0,0), whose logging pixel they never set.
__global__ void k_ray_count_origin(const unsigned long long* rays, uint32_t* rayCount) {
  if (threadIdx.x == 0 && blockIdx.x == 0) rayCount[0] += (uint32_t)(rays[0] + rays[1] - rays[2]);
}

// ------------------------------------------------------------------ batched frames (device)
//
// A batched frame renders many contexts (scenes, cameras, sizes, options) with one launch
// per stage: the per-pixel kernels' blocks cover the tiles of every context, the persistent
// task kernels drain every context's task lists in turn, and the scans run over every
// context's lists at once.  What differs between contexts and frames (arguments, buffers,
// which stages run) is read from the batch table, which the host rewrites before each frame,
// so the launches themselves (a graph) stay the same from frame to frame.
//
// The table is read through a constant-address-space pointer: a wave-uniform context index
// then turns every argument load into a scalar load, as kernel arguments are.

// Stages a context runs in this frame.
enum BatchStage : uint32_t {
  kBRobust = 1u << 0,     // RobustReuseOptimization (Robust gather)
  kBCollect = 1u << 1,    // Collect
  kBSplat = 1u << 2,      // Reproject or MultiReproject
  kBMulti = 1u << 3,      // MultiReproject
  kBTasks = 1u << 4,      // splat sort and candidate-compacted temporal resampling
  kBGather = 1u << 5,     // Gather
  kBNeighbors = 1u << 6,  // spatial reuse has neighbours (iterations run: spatialIters)
};

// The batch's segmented scans: task lists (flags -> queue positions, live entries enqueued)
// and splat cells (counts -> offsets).
enum BatchScanKind : uint32_t { kScanRobust = 0, kScanTemporal = 1, kScanSpatial = 2, kScanCells = 3, kBatchScans = 4 };
static constexpr uint32_t kBatchScanChunk = 2048;  // entries per block (256 lanes x 8)

struct BatchScan {
  const uint32_t* in;  // flags or counts
  uint32_t* out;       // exclusive scan of `in`
  uint32_t* queue;     // task lists: the pixel of every live entry; null for counts
  uint32_t* fetch;     // task lists: the persistent kernel's next task
  uint32_t n;          // entries this frame (0: the context does not run this scan)
  uint32_t stride;     // task lists: tile-order positions per list
  uint32_t chunk0;     // first chunk of the segment in the kind's chunk space
  uint32_t chunks;     // chunks reserved for the segment
};

// Copies of a frame: [0] the caller's motion vectors in, [1] and [2] the initial candidates
// to the capture slot, [3] the colour out.  Sizes in 8-byte units (0: no copy).
static constexpr uint32_t kBatchCopies = 4;

struct BatchCtx {
  uint32_t stages;       // BatchStage bits
  uint32_t spatialIters; // spatial iterations this frame
  uint32_t args;         // KernelArgs of the temporal phase; spatial iteration i at args + 1 + i
  uint32_t tile0;        // first tile of the context in the batch's tile space
  uint32_t tiles;        // tiles of the frame
  uint32_t gx;           // tiles per row
  uint32_t kind;         // TemporalTaskKind
  uint32_t P, KPP;
  TaskLists tl;
  SpatialTasks sp;
  TaskQueue rq;
  uint32_t* largeCells;  // [0] count, then the cells too long for a thread to sort
  const PackedReservoir* finalR;
  BatchScan scan[kBatchScans];
  uint2* copyDst[kBatchCopies];
  const uint2* copySrc[kBatchCopies];
  uint32_t copyN[kBatchCopies];
  uint32_t* rayCount;    // caller's per-pixel ray counts, or null
  uint32_t* pathLength;  // caller's per-pixel path lengths, or null
};

static constexpr uint32_t kBatchHeaderBytes = 256;
struct BatchHeader {
  uint32_t n;              // contexts
  uint32_t argsOffset;     // byte offset of the KernelArgs array
  uint32_t* chunkSums;     // chunk sums of every scan kind
  uint32_t chunkBase[kBatchScans];
};
static_assert(sizeof(BatchHeader) <= kBatchHeaderBytes, "batch header");

typedef const __attribute__((address_space(4))) unsigned char* BatchTab;

__device__ __forceinline__ const BatchHeader& bt_head(BatchTab t) { return *(const BatchHeader*)t; }
__device__ __forceinline__ const BatchCtx& bt_ctx(BatchTab t, uint32_t c) {
  return ((const BatchCtx*)(t + kBatchHeaderBytes))[c];
}
__device__ __forceinline__ const KernelArgs& bt_args(BatchTab t, uint32_t i) {
  return ((const KernelArgs*)(t + bt_head(t).argsOffset))[i];
}

// Context and tile of global tile g among contexts c0 .. c1-1 (consecutive tile ranges).
__device__ __forceinline__ bool batch_tile(BatchTab t, uint32_t c0, uint32_t c1, uint32_t g, uint32_t& c, TileId& tile) {
  c = c0;
  while (c + 1 < c1 && g >= bt_ctx(t, c + 1).tile0) c++;
  c = __builtin_amdgcn_readfirstlane(c);
  const BatchCtx& bc = bt_ctx(t, c);
  const uint32_t l = g - bc.tile0;
  tile = TileId{l % bc.gx, l / bc.gx, bc.gx};
  return l < bc.tiles;
}

// Per-pixel batched kernels: blocks cover tiles t0 .. t0+gridDim.x-1 of contexts c0 .. c1-1.
#define RS_BATCH_PIXEL(cond)                                         \
  uint32_t ci;                                                       \
  TileId tile;                                                       \
  if (!batch_tile(t, c0, c1, t0 + blockIdx.x, ci, tile)) return;     \
  const BatchCtx& bc = bt_ctx(t, ci);                                \
  if (!(cond)) return;                                               \
  const KernelArgs& a = bt_args(t, bc.args);

__global__ __launch_bounds__(256) void kb_vbuffer(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(true)
  vbuffer_body(a, tile);
}

template <bool kEnv, bool kStats>
RS_HEAVY_KERNEL(RS_WAVES_INITIAL) void kb_initial(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(true)
  initial_body<kEnv, kStats>(a, tile);
}

__global__ __launch_bounds__(256) void kb_robust_flags(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBRobust)
  robust_flags_body(a, tile, bc.rq);
}

__global__ __launch_bounds__(256) void kb_collect(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBCollect)
  collect_body(a, tile);
}

__global__ __launch_bounds__(256) void kb_reproject(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL((bc.stages & (kBSplat | kBMulti)) == kBSplat)
  reproject_body(a, tile);
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void kb_multi_reproject(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBMulti)
  multi_reproject_body<kEnv>(a, tile);
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void kb_gather(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBGather)
  gather_body<kEnv>(a, tile);
}

__global__ __launch_bounds__(256) void kb_temporal_flags(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBTasks)
  temporal_flags_body(a, tile, bc.tl, bc.kind == kTasksBackup);
}

__global__ __launch_bounds__(256) void kb_temporal_combine(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(bc.stages & kBTasks)
  temporal_combine_body(a, tile, bc.tl, bc.kind);
}

__global__ __launch_bounds__(256) void kb_spatial_flags(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0, uint32_t it) {
  RS_BATCH_PIXEL((bc.stages & kBNeighbors) && it < bc.spatialIters)
  spatial_flags_body(bt_args(t, bc.args + 1 + it), tile, bc.sp);
}

__global__ __launch_bounds__(256) void kb_spatial_combine(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0, uint32_t it) {
  RS_BATCH_PIXEL(it < bc.spatialIters)
  spatial_combine_body(bt_args(t, bc.args + 1 + it), tile, bc.sp);
}

__global__ __launch_bounds__(256) void kb_resolve(BatchTab t, uint32_t c0, uint32_t c1, uint32_t t0) {
  RS_BATCH_PIXEL(true)
  uint32_t x, y;
  if (lane_pixel(a.p, tile, x, y)) resolve(a.p, a.b, bc.finalR, x, y);
}

// True when the list behind `fetch` is already exhausted (a plain read, so that waves
// passing over finished contexts do not queue up on their counters).
__device__ __forceinline__ bool list_drained(uint32_t* fetch, uint32_t total) {
  return __builtin_amdgcn_readfirstlane(__hip_atomic_load(fetch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >= total;
}

// Persistent batched task kernels: every wave drains the lists of contexts c0 .. c1-1 in
// turn, adding its rays to each context's counters as it leaves it.
template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void kb_robust_tasks(BatchTab t, uint32_t c0, uint32_t c1) {
  RS_TASK_PROLOGUE
  for (uint32_t ci = c0; ci < c1; ci++) {
    const BatchCtx& bc = bt_ctx(t, ci);
    if (!(bc.stages & kBRobust)) continue;
    const KernelArgs& a = bt_args(t, bc.args);
    if (list_drained(bc.rq.fetch, 8u * queue_start(bc.rq, 1))) continue;
    tr.bvh = a.bvh;
    DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
    robust_tasks_run(c, a, bc.rq);
    RS_TRACE_EPILOGUE
    tr.rays_closest = 0;
    tr.rays_visibility = 0;
  }
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_TEMPORAL) void kb_temporal_tasks(BatchTab t, uint32_t c0, uint32_t c1) {
  RS_TASK_PROLOGUE
  for (uint32_t ci = c0; ci < c1; ci++) {
    const BatchCtx& bc = bt_ctx(t, ci);
    if (!(bc.stages & kBTasks)) continue;
    const KernelArgs& a = bt_args(t, bc.args);
    if (list_drained(bc.tl.q.fetch, queue_start(bc.tl.q, bc.tl.q.lists) + splat_total(a.b, bc.KPP))) continue;
    tr.bvh = a.bvh;
    DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
    temporal_tasks_run(c, a, bc.tl, bc.kind);
    RS_TRACE_EPILOGUE
    tr.rays_closest = 0;
    tr.rays_visibility = 0;
  }
}

template <bool kEnv>
RS_HEAVY_KERNEL(RS_WAVES_SPATIAL) void kb_spatial_tasks(BatchTab t, uint32_t c0, uint32_t c1, uint32_t it) {
  RS_TASK_PROLOGUE
  for (uint32_t ci = c0; ci < c1; ci++) {
    const BatchCtx& bc = bt_ctx(t, ci);
    if (!(bc.stages & kBNeighbors) || it >= bc.spatialIters) continue;
    const KernelArgs& a = bt_args(t, bc.args + 1 + it);
    if (list_drained(bc.sp.q.fetch, queue_start(bc.sp.q, bc.sp.q.lists))) continue;
    tr.bvh = a.bvh;
    DevCtx c{&a.p, &a.s, &tr, kEnv && a.s.env.present != 0u};
    spatial_tasks_run(c, a, bc.sp);
    RS_TRACE_EPILOGUE
    tr.rays_closest = 0;
    tr.rays_visibility = 0;
  }
}

// Per-context work with a grid of (blocks, contexts): blockIdx.y is the context.

// Frame start: the ray counters, and for contexts that splat this frame the splat cells and
// the long-cell list.
__global__ void kb_begin(BatchTab t) {
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  const KernelArgs& a = bt_args(t, bc.args);
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x, step = gridDim.x * blockDim.x;
  if (i0 < 8) reinterpret_cast<uint32_t*>(a.rays)[i0] = 0u;
  if (!(bc.stages & kBSplat)) return;
  if (i0 == 0 && (bc.stages & kBTasks)) bc.largeCells[0] = 0u;
  for (uint32_t i = i0; i < bc.KPP; i += step) a.b.cellCount[i] = 0u;
}

// Copies [first, first + count) of every context.
__global__ void kb_copy(BatchTab t, uint32_t first, uint32_t count) {
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x, step = gridDim.x * blockDim.x;
  for (uint32_t k = first; k < first + count; k++) {
    const uint32_t n = bc.copyN[k];
    uint2* d = bc.copyDst[k];
    const uint2* s = bc.copySrc[k];
    for (uint32_t i = i0; i < n; i += step) d[i] = s[i];
  }
}

__global__ void kb_splat_fill(BatchTab t) {
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  if (!(bc.stages & kBTasks)) return;
  const KernelArgs& a = bt_args(t, bc.args);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < bc.KPP; i += gridDim.x * blockDim.x)
    splat_fill_entry(a.b, bc.tl.sortedCell, i, bc.P);
}

__global__ void kb_splat_order(BatchTab t) {
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  if (!(bc.stages & kBTasks)) return;
  const KernelArgs& a = bt_args(t, bc.args);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < bc.KPP; i += gridDim.x * blockDim.x)
    splat_order_cell(a.b, i, bc.largeCells + 1, bc.largeCells);
}

__global__ __launch_bounds__(256) void kb_splat_order_large(BatchTab t) {
  __shared__ uint32_t s[kSortChunk];
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  if (!(bc.stages & kBTasks)) return;
  const KernelArgs& a = bt_args(t, bc.args);
  const uint32_t nLarge = bc.largeCells[0];
  for (uint32_t li = blockIdx.x; li < nLarge; li += gridDim.x)
    splat_order_large_cell(a.b, bc.largeCells[1 + li], a.b.splatRank, s);
}

// Frame end: the resampling kernels' rays at pixel (0,0) and the cleared path lengths.
__global__ void kb_end(BatchTab t) {
  const BatchCtx& bc = bt_ctx(t, blockIdx.y);
  const KernelArgs& a = bt_args(t, bc.args);
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (bc.rayCount && i0 == 0) bc.rayCount[0] += (uint32_t)(a.rays[0] + a.rays[1] - a.rays[2]);
  if (bc.pathLength)
    for (uint32_t i = i0; i < bc.P; i += gridDim.x * blockDim.x) bc.pathLength[i] = 0u;
}

// Exclusive scan of one value per lane over the 256-lane block; `total` receives the sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t s_wave[4];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t incl = v;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63u) s_wave[w] = incl;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t k = 0; k < w; k++) base += s_wave[k];
  total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
  __syncthreads();
  return base + incl - v;
}

// Segment of chunk g of a scan kind among the contexts.
__device__ __forceinline__ uint32_t scan_segment(BatchTab t, uint32_t kind, uint32_t g) {
  const uint32_t n = bt_head(t).n;
  uint32_t c = 0;
  while (c + 1 < n && g >= bt_ctx(t, c + 1).scan[kind].chunk0) c++;
  return __builtin_amdgcn_readfirstlane(c);
}

// The lane's 8 consecutive entries of a chunk, zero past the segment's end.
__device__ __forceinline__ void scan_load8(const uint32_t* in, uint32_t i0, uint32_t n, uint32_t v[8]) {
  if (i0 + 8 <= n) {
    const uint4 p = *reinterpret_cast<const uint4*>(in + i0), q = *reinterpret_cast<const uint4*>(in + i0 + 4);
    v[0] = p.x; v[1] = p.y; v[2] = p.z; v[3] = p.w;
    v[4] = q.x; v[5] = q.y; v[6] = q.z; v[7] = q.w;
  } else {
    for (uint32_t k = 0; k < 8; k++) v[k] = i0 + k < n ? in[i0 + k] : 0u;
  }
}

// Pass 1: the sum of every chunk.
__global__ __launch_bounds__(256) void kb_scan_reduce(BatchTab t, uint32_t kind) {
  const uint32_t g = blockIdx.x, c = scan_segment(t, kind, g);
  const BatchScan& sc = bt_ctx(t, c).scan[kind];
  const uint32_t lo = (g - sc.chunk0) * kBatchScanChunk;
  if (g - sc.chunk0 >= sc.chunks || lo >= sc.n) return;
  uint32_t v[8];
  scan_load8(sc.in, lo + threadIdx.x * 8u, sc.n, v);
  uint32_t sum = 0;
  for (uint32_t k = 0; k < 8; k++) sum += v[k];
  uint32_t total;
  block_exclusive_scan(sum, total);
  if (threadIdx.x == 0) bt_head(t).chunkSums[bt_head(t).chunkBase[kind] + g] = total;
}

// Pass 2, one block per context: the chunk sums of its segment, exclusively scanned in place.
__global__ __launch_bounds__(256) void kb_scan_chunks(BatchTab t, uint32_t kind) {
  const BatchScan& sc = bt_ctx(t, blockIdx.x).scan[kind];
  if (sc.n == 0) return;
  uint32_t* sums = bt_head(t).chunkSums + bt_head(t).chunkBase[kind] + sc.chunk0;
  const uint32_t m = (sc.n + kBatchScanChunk - 1) / kBatchScanChunk;
  uint32_t carry = 0;
  for (uint32_t k0 = 0; k0 < m; k0 += 256) {
    const uint32_t k = k0 + threadIdx.x;
    const uint32_t v = k < m ? sums[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, total);
    if (k < m) sums[k] = carry + ex;
    carry += total;
  }
}

// Pass 3: every entry's exclusive scan; for task lists, every live entry's pixel is written
// at its queue position and the fetch counter is reset.
__global__ __launch_bounds__(256) void kb_scan_apply(BatchTab t, uint32_t kind) {
  const uint32_t g = blockIdx.x, c = scan_segment(t, kind, g);
  const BatchCtx& bc = bt_ctx(t, c);
  const BatchScan& sc = bc.scan[kind];
  const uint32_t lo = (g - sc.chunk0) * kBatchScanChunk;
  if (g - sc.chunk0 >= sc.chunks || lo >= sc.n) return;
  if (sc.queue && g == sc.chunk0 && threadIdx.x == 0) *sc.fetch = 0u;
  const uint32_t i0 = lo + threadIdx.x * 8u;
  uint32_t v[8];
  scan_load8(sc.in, i0, sc.n, v);
  uint32_t sum = 0;
  for (uint32_t k = 0; k < 8; k++) sum += v[k];
  uint32_t total;
  uint32_t run = bt_head(t).chunkSums[bt_head(t).chunkBase[kind] + g] + block_exclusive_scan(sum, total);
  const FrameParams& p = bt_args(t, bc.args).p;
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t i = i0 + k;
    if (i >= sc.n) break;
    sc.out[i] = run;
    if (sc.queue && v[k]) {
      const uint32_t pos = i % sc.stride, tl = pos >> 8;
      uint32_t x, y;
      tile_lane_pixel(p, TileId{tl % bc.gx, tl / bc.gx, bc.gx}, pos & 255u, x, y);
      sc.queue[run] = y * p.W + x;
    }
    run += v[k];
  }
}

// The task lists of single frames are scanned the same way, reduce-then-scan in chunks of
// kBatchScanChunk entries: no block waits on another's result, so the scan keeps its speed
// while a persistent kernel on the side stream holds most of the device.  (A decoupled
// look-back scan, whose blocks wait on their predecessors, took 0.40 ms there instead of
// 0.02 ms.)
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* in, uint32_t n, uint32_t* sums) {
  uint32_t v[8];
  scan_load8(in, blockIdx.x * kBatchScanChunk + threadIdx.x * 8u, n, v);
  uint32_t sum = 0;
  for (uint32_t k = 0; k < 8; k++) sum += v[k];
  uint32_t total;
  block_exclusive_scan(sum, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: the m chunk sums, exclusively scanned in place.
__global__ __launch_bounds__(256) void k_scan_chunks(uint32_t* sums, uint32_t m) {
  uint32_t carry = 0;
  for (uint32_t k0 = 0; k0 < m; k0 += 256) {
    const uint32_t k = k0 + threadIdx.x;
    const uint32_t v = k < m ? sums[k] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, total);
    if (k < m) sums[k] = carry + ex;
    carry += total;
  }
}

// Every flag's exclusive scan into q.offsets, every live entry's pixel at its queue position
// (tile-order position i % stride of a grid gx tiles wide), and the fetch counter reset.
__global__ __launch_bounds__(256) void k_scan_queue(FrameParams p, TaskQueue q, const uint32_t* sums, uint32_t gx) {
  const uint32_t n = q.lists * q.stride;
  if (blockIdx.x == 0 && threadIdx.x == 0) *q.fetch = 0u;
  const uint32_t i0 = blockIdx.x * kBatchScanChunk + threadIdx.x * 8u;
  uint32_t v[8];
  scan_load8(q.flags, i0, n, v);
  uint32_t sum = 0;
  for (uint32_t k = 0; k < 8; k++) sum += v[k];
  uint32_t total;
  uint32_t run = sums[blockIdx.x] + block_exclusive_scan(sum, total);
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t i = i0 + k;
    if (i >= n) break;
    q.offsets[i] = run;
    if (v[k]) {
      const uint32_t pos = i % q.stride, tl = pos >> 8;
      uint32_t x, y;
      tile_lane_pixel(p, TileId{tl % gx, tl / gx, gx}, pos & 255u, x, y);
      q.queue[run] = y * p.W + x;
    }
    run += v[k];
  }
}

}  // namespace

// ------------------------------------------------------------------ host context

namespace {

// Tree builder: parallel locally-ordered clustering (default) or the plain LBVH.
#ifndef RS_BVH_LBVH
static constexpr bool kUseClustering = true;
#else
static constexpr bool kUseClustering = false;
#endif
// The clustering build collapses to 4-wide nodes by the surface-area dynamic program;
// RS_BVH4_GREEDY=1 selects the greedy largest-child collapse instead (always used by LBVH).
#ifndef RS_BVH4_GREEDY
#define RS_BVH4_GREEDY 0
#endif
static constexpr size_t kMaxClusterIterations = 1u << 16;
static constexpr uint32_t kMaxBatchQueues = 8;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

// Device buffers of one compacted task list (TaskQueue).
struct QueueBufs {
  DevBuf flags, offsets, queue, fetch, chunkSums;
};

struct rs_ctx {
  rs_options opt;
  std::string err;
  int device = 0;

  // scene
  bool hasScene = false;
  uint32_t numTris = 0, numVerts = 0, numMaterials = 0, numLights = 0;
  uint32_t diffuseModel = 0;
  float invWeightsSum = 0.f;
  DevBuf positions, normals, indices, triMaterial, triToLight, materials, lights, alias;
  std::vector<uint32_t> hIndices;
  // bvh
  DevBuf bvhNodes, bvhTris, bvhScratch, bvh4Nodes;
  DevBuf shifted;  // Robust gather: 8 shifted paths per pixel
  DevBuf scatterCounts;
  bool splatValid = false;  // the last frame produced splat lists
  // frozen replay (adjoint)
  FrameParams lastParams{};
  bool rendered = false;
  DevBuf frozen, frozenC, frozenParams0;
  uint32_t frozenW = 0, frozenH = 0;
  uint32_t bvhRoot = 0;
  uint32_t bvh4Count = 0, bvh4Levels = 0, buildIterations = 0;
  uint32_t bvhBuild = RS_BVH_BUILD_CLUSTERING;
  float sceneScale = 0.f;  // largest vertex coordinate magnitude (the walks' culling slack)
  DevBuf scaleBits;
  DevBuf bvh4Src;                                          // clustering node behind each child slot
  std::vector<std::pair<uint32_t, uint32_t>> plocIters;    // (first node id, count) per iteration
  PlocBuild ploc{};                                        // buffers of the last clustering build
  bool refitReady = false;
  // environment map (device texels and importance mips; host-side parameters)
  DevBuf envTexels, envImportance;
  EnvMapView env{};
  std::vector<uint32_t> hTriMaterial;
  std::vector<rs_material> hMaterials;
  // instances: triangle range per instance (x first, y count), the object-space vertices,
  // the instance of each vertex (kNoInstance outside every instance) and the transforms
  std::vector<rs_instance_desc> instances;
  DevBuf instTri, objPositions, objNormals, vertInstance, instXforms, lightFlux;
  uint32_t frozenMaterials = 0;  // material count at rs_freeze
  FrameParams frozenFrame{};     // the frozen frame's parameters (camera, seeds, options)

  // frame state
  uint32_t W = 0, H = 0;
  uint32_t frameCount = 0, seed = 0;
  bool resetPending = true;
  bool sceneFrozen = false;  // Scene::isFrozen as the pass sees it
  // While set, a buffer that would have to grow is an error instead of an allocation (a
  // batch frame being recorded into a graph).
  bool noAlloc = false;
  rs_frame_flags flags{0, 1, 0};
  CameraManager cm{};
  uint32_t prevNumKeyframes = 0;
  rs_camera lastCam{};

  // per-pixel buffers: reservoir/recon slots 0..4 (curr, prev, backup, capture-initial,
  // the frozen scene's spatial ping-pong); which slot holds what moves with the swaps
  DevBuf R[5], C[5];
  int curr = 0, prev = 1, tmp = 4;
  int lastFinal = 1, lastTemporal = 0;  // slots of the last frame's final and temporal output
  bool temporalValid = false;
  uint32_t splatPartitions = 1;  // K of the splat lists of the last frame
  uint32_t tileShift = 4;        // 16-pixel-wide tiles
  DevBuf mvec, vbuffer, floatingCoords, splatDst, cellCount, cellOffset, splatRank, sorted, color, neighborOffsets,
      rays, scanTemp;
  size_t scanTempBytes = 0;
  // candidate-compacted resampling: candidate slots, the cell of each splat-list entry,
  // the task lists and the persistent task kernels' grids (per kernel instantiation:
  // [0] without, [1] with an environment map)
  DevBuf cand, sortedCell, spatialM1, spatialCand, largeCells, splatFetch;
  QueueBufs temporalQ, spatialQ, robustQ, multiQ;
  uint32_t taskBlocks[2] = {0, 0}, spatialBlocks[2] = {0, 0}, robustBlocks[2] = {0, 0}, multiBlocks[2] = {0, 0};
  // stages overlapped with the initial candidates: the side stream and its fork/join events
  bool overlap = true;
  hipStream_t side = nullptr;
  hipEvent_t forkEv = nullptr, joinEv = nullptr, splatEv = nullptr;

  // timing
  bool timing = false;
  hipEvent_t ev[16] = {};
  rs_stats stats{};
};

namespace {

int32_t fail(rs_ctx* c, int32_t code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(call)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, RS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int32_t ensure(rs_ctx* ctx, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && (b.ptr || bytes == 0)) return RS_OK;
  if (ctx->noAlloc)
    return fail(ctx, RS_E_SIZE_MISMATCH,
                "a buffer must grow, which this call may not allocate: render the frame once with rs_render_frame");
  if (b.ptr) hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  if (bytes == 0) return RS_OK;
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) return fail(ctx, RS_E_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  b.bytes = bytes;
  return RS_OK;
}

void release(DevBuf& b) {
  if (b.ptr) hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

// Neighbor-offset table: R2 sequence in the unit disk, RG8Snorm, identical generator to
// RS/ReservoirSplatting.cpp:1072-1097.
void make_neighbor_offsets(std::vector<int8_t>& out, uint32_t count) {
  out.assign(2 * count, 0);
  const int R = 254;
  const float phi2 = 1.f / 1.3247179572447f;
  float u = 0.5f, v = 0.5f;
  for (uint32_t index = 0; index < count * 2;) {
    u += phi2;
    v += phi2 * phi2;
    if (u >= 1.f) u -= 1.f;
    if (v >= 1.f) v -= 1.f;
    float rSq = (u - 0.5f) * (u - 0.5f) + (v - 0.5f) * (v - 0.5f);
    if (rSq > 0.25f) continue;
    out[index++] = (int8_t)((u - 0.5f) * R);
    out[index++] = (int8_t)((v - 0.5f) * R);
  }
}

int32_t validate(rs_options& o, std::string& msg) {
  bool clamped = false;
  if (o.quirkMask & ~RS_QUIRKS_REFERENCE) {
    o.quirkMask &= RS_QUIRKS_REFERENCE;
    msg = "'quirkMask' has unknown bits. Keeping the known RS_QUIRK_* bits.";
    clamped = true;
  }
  if (o.specularRoughnessThreshold < 0.f || o.specularRoughnessThreshold > 1.f) {
    o.specularRoughnessThreshold = std::min(std::max(o.specularRoughnessThreshold, 0.f), 1.f);
    msg = "'specularRoughnessThreshold' has invalid value. Clamping to range [0,1].";
    clamped = true;
  }
  if (o.samplesPerPixel < 1 || o.samplesPerPixel > 16) {
    o.samplesPerPixel = std::min(std::max(o.samplesPerPixel, 1u), 16u);
    msg = "'samplesPerPixel' must be in the range [1, 16]. Clamping to this range.";
    clamped = true;
  }
  auto cb = [&](uint32_t& v, const char* name) {
    if (v > 254) {
      v = 254;
      msg = std::string("'") + name + "' exceeds the maximum supported bounces. Clamping to 254.";
      clamped = true;
    }
  };
  cb(o.maxSurfaceBounces, "maxSurfaceBounces");
  cb(o.maxDiffuseBounces, "maxDiffuseBounces");
  cb(o.maxSpecularBounces, "maxSpecularBounces");
  cb(o.maxTransmissionBounces, "maxTransmissionBounces");
  uint32_t minSurface = std::max(o.maxDiffuseBounces, std::max(o.maxSpecularBounces, o.maxTransmissionBounces));
  o.maxSurfaceBounces = std::max(o.maxSurfaceBounces, minSurface);
  if (o.temporalReuse > 3 || o.gatherMechanism > 2 || o.scatterBackupMIS > 1 || o.misHeuristic > 2) {
    msg = "enum option out of range";
    return RS_E_INVALID_ARG;
  }
  if (o.numTimePartitions < 1) {
    o.numTimePartitions = 1;
    msg = "'numTimePartitions' must be at least 1.";
    clamped = true;
  }
  if (o.numTimePartitions > kMaxTimePartitions) {
    o.numTimePartitions = kMaxTimePartitions;
    msg = "'numTimePartitions' exceeds the supported maximum. Clamping to 64.";
    clamped = true;
  }
  if (!(o.artificialFrameTime > 0.f)) {
    msg = "'artificialFrameTime' must be positive";
    return RS_E_INVALID_ARG;
  }
  return clamped ? RS_E_BAD_OPTION : RS_OK;
}

// Deterministic Vose alias table (no shuffle), packed like
// Source/Falcor/Rendering/Lights/EmissivePowerSampler.cpp:157-164.
void build_alias_table(const std::vector<float>& weightsIn, std::vector<uvec2>& table, double& sumOut) {
  uint32_t N = (uint32_t)weightsIn.size();
  std::vector<float> weights = weightsIn;
  double sum = 0.0;
  for (float f : weights) sum += f;
  for (float& f : weights) f *= (float)N / (float)sum;
  std::vector<uint32_t> perm(N);
  for (uint32_t i = 0; i < N; i++) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return weights[a] < weights[b]; });
  std::vector<float> thresholds(N, 1.f);
  std::vector<uint32_t> redirect(N);
  for (uint32_t i = 0; i < N; i++) redirect[i] = i;
  if (N > 1) {
    uint32_t head = 0, tail = N - 1;
    while (head != tail) {
      uint32_t i = perm[head], j = perm[tail];
      thresholds[i] = weights[i];
      redirect[i] = j;
      weights[j] -= 1.f - weights[i];
      if (head == tail - 1) {
        thresholds[j] = 1.f;
        redirect[j] = j;
        break;
      } else if (weights[j] < 1.f) {
        std::swap(perm[head], perm[tail]);
        tail--;
      } else {
        head++;
      }
    }
  }
  table.resize(N);
  for (uint32_t i = 0; i < N; i++) {
    uint32_t prob = f32tof16(thresholds[i]) << 16u;
    uint32_t lx = redirect[i] & 0xFFFFFFu, ly = i & 0xFFFFFFu;
    table[i] = uvec2{prob | ((lx >> 8u) & 0xFFFFu), ((lx & 0xFFu) << 24u) | ly};
  }
  sumOut = sum;
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

void rs_default_options(rs_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->samplesPerPixel = 1;
  o->maxSurfaceBounces = 9;
  o->maxDiffuseBounces = 9;
  o->maxSpecularBounces = 9;
  o->maxTransmissionBounces = 9;
  o->useRussianRoulette = 1;
  o->misHeuristic = RS_MIS_BALANCE;
  o->misPowerExponent = 2.f;
  o->specularRoughnessThreshold = 0.25f;
  o->adjustShadingNormals = 0;
  o->gbufferAdjustShadingNormals = 1;
  o->enableTemporalResampling = 1;
  o->temporalReuse = RS_TEMPORAL_GATHER_ONLY;
  o->gatherMechanism = RS_GATHER_FAST;
  o->scatterBackupMIS = RS_BACKUP_MIS_BALANCE;
  o->numTimePartitions = 2;
  o->useConfidenceWeightsTemporally = 1;
  o->enableSpatialResampling = 1;
  o->numSpatialIterations = 1;
  o->neighborCount = 1;
  o->useConfidenceWeightsSpatially = 1;
  o->spatialGatherRadius = 30.f;
  o->artificialFrameTime = 1.f / 60.f;
  o->useFixedSeed = 0;
  o->fixedSeed = 1;
  o->captureStages = 0;
  o->quirkMask = RS_QUIRKS_REFERENCE;
}

int32_t rs_create(const rs_options* opts, rs_ctx** out) {
  if (!out) return RS_E_INVALID_ARG;
  *out = nullptr;
  rs_ctx* ctx = new rs_ctx();
  if (const char* e = getenv("RS_STAGE_OVERLAP")) ctx->overlap = strtol(e, nullptr, 10) != 0;
  if (opts) ctx->opt = *opts;
  else rs_default_options(&ctx->opt);
  std::string msg;
  int32_t v = validate(ctx->opt, msg);
  if (v == RS_E_INVALID_ARG) {
    delete ctx;
    return RS_E_INVALID_ARG;
  }
  *out = ctx;
  if (v == RS_E_BAD_OPTION) {
    ctx->err = msg;
    return RS_E_BAD_OPTION;
  }
  return RS_OK;
}

void rs_destroy(rs_ctx* ctx) {
  if (!ctx) return;
  if (ctx->neighborOffsets.ptr) (void)hipDeviceSynchronize();
  DevBuf* all[] = {&ctx->positions, &ctx->normals, &ctx->indices, &ctx->triMaterial, &ctx->triToLight,
                   &ctx->materials, &ctx->lights, &ctx->alias, &ctx->bvhNodes, &ctx->bvhTris, &ctx->bvhScratch,
                   &ctx->mvec, &ctx->vbuffer, &ctx->floatingCoords, &ctx->splatDst, &ctx->cellCount,
                   &ctx->cellOffset, &ctx->splatRank, &ctx->sorted, &ctx->color, &ctx->neighborOffsets, &ctx->rays,
                   &ctx->scanTemp, &ctx->bvh4Nodes, &ctx->shifted, &ctx->scatterCounts,
                   &ctx->bvh4Src, &ctx->frozen, &ctx->frozenC, &ctx->frozenParams0, &ctx->cand,
                   &ctx->sortedCell, &ctx->spatialM1, &ctx->spatialCand, &ctx->largeCells, &ctx->splatFetch, &ctx->temporalQ.flags,
                   &ctx->temporalQ.offsets, &ctx->temporalQ.queue, &ctx->temporalQ.fetch, &ctx->temporalQ.chunkSums,
                   &ctx->spatialQ.flags, &ctx->spatialQ.offsets, &ctx->spatialQ.queue, &ctx->spatialQ.fetch,
                   &ctx->spatialQ.chunkSums, &ctx->robustQ.flags, &ctx->robustQ.offsets, &ctx->robustQ.queue,
                   &ctx->robustQ.fetch, &ctx->robustQ.chunkSums, &ctx->multiQ.flags, &ctx->multiQ.offsets,
                   &ctx->multiQ.queue, &ctx->multiQ.fetch, &ctx->multiQ.chunkSums, &ctx->scaleBits, &ctx->envTexels,
                   &ctx->envImportance, &ctx->instTri,
                   &ctx->objPositions, &ctx->objNormals, &ctx->vertInstance, &ctx->instXforms, &ctx->lightFlux};
  for (DevBuf* b : all) release(*b);
  for (int i = 0; i < 5; i++) {
    release(ctx->R[i]);
    release(ctx->C[i]);
  }
  for (auto& e : ctx->ev)
    if (e) hipEventDestroy(e);
  if (ctx->forkEv) hipEventDestroy(ctx->forkEv);
  if (ctx->joinEv) hipEventDestroy(ctx->joinEv);
  if (ctx->splatEv) hipEventDestroy(ctx->splatEv);
  if (ctx->side) hipStreamDestroy(ctx->side);
  delete ctx;
}

const char* rs_last_error(const rs_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Device-side state every frame needs (neighbor-offset table, ray counters); created on
// first use so that a context can be created and configured without a GPU.
static int32_t ensure_device(rs_ctx* ctx) {
  if (ctx->neighborOffsets.ptr) return RS_OK;
  HIP_TRY(hipGetDevice(&ctx->device));
  // created here, with the context's first buffers, so that a frame creates nothing
  if (!ctx->side) HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  if (!ctx->forkEv) HIP_TRY(hipEventCreateWithFlags(&ctx->forkEv, hipEventDisableTiming));
  if (!ctx->joinEv) HIP_TRY(hipEventCreateWithFlags(&ctx->joinEv, hipEventDisableTiming));
  if (!ctx->splatEv) HIP_TRY(hipEventCreateWithFlags(&ctx->splatEv, hipEventDisableTiming));
  std::vector<int8_t> offs;
  make_neighbor_offsets(offs, kSpatialNeighborSamples);
  // rays: [0] closest-hit and [1] visibility queries of the frame, [2] those of the
  // initial-candidate kernel, [3] unused
  if (ensure(ctx, ctx->neighborOffsets, offs.size()) || ensure(ctx, ctx->rays, 4 * sizeof(unsigned long long)))
    return RS_E_OUT_OF_MEMORY;
  HIP_TRY(hipMemcpy(ctx->neighborOffsets.ptr, offs.data(), offs.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(ctx->rays.ptr, 0, 4 * sizeof(unsigned long long)));
  return RS_OK;
}

int32_t rs_build_alias_table(const float* weights, uint32_t n, uint32_t* table, float* invWeightsSum) {
  if (!weights || !table || !invWeightsSum || n == 0 || n > (1u << 24)) return RS_E_INVALID_ARG;
  std::vector<float> w(weights, weights + n);
  double sum = 0.0;
  for (float f : w)
    if (!(f >= 0.f)) return RS_E_INVALID_ARG;
  std::vector<uvec2> t;
  build_alias_table(w, t, sum);
  if (!(sum > 0.0)) return RS_E_INVALID_ARG;
  std::memcpy(table, t.data(), 8 * (size_t)n);
  *invWeightsSum = (float)(1.0 / sum);
  return RS_OK;
}

int32_t rs_set_options(rs_ctx* ctx, const rs_options* opts) {
  if (!ctx || !opts) return RS_E_INVALID_ARG;
  rs_options o = *opts;
  std::string msg;
  int32_t v = validate(o, msg);
  if (v == RS_E_INVALID_ARG) return fail(ctx, v, msg);
  ctx->opt = o;
  if (v == RS_E_BAD_OPTION) return fail(ctx, v, msg);
  return RS_OK;
}

int32_t rs_get_options(const rs_ctx* ctx, rs_options* out) {
  if (!ctx || !out) return RS_E_INVALID_ARG;
  *out = ctx->opt;
  return RS_OK;
}

// Deepest 4-wide tree a walk can traverse: a walk pushes at most three children per level,
// so three entries per level must fit the traversal stack.  RS_BVH_MAX_LEVELS in the
// environment lowers the limit (to exercise the fallback and the failure path).
static uint32_t bvh_level_limit() {
  uint32_t lim = (uint32_t)kStackMax / 3;
  if (const char* e = getenv("RS_BVH_MAX_LEVELS")) {
    const long v = strtol(e, nullptr, 10);
    if (v > 0 && (unsigned long)v < lim) lim = (uint32_t)v;
  }
  return lim;
}

// Balanced binary hierarchy over n >= 2 leaves in Morton order, in the LBVH layout: the node
// over leaves [lo, hi] splits after leaf lo + (hi - lo) / 2, so the depth is ceil(log2 n).
// Internal nodes are numbered as the Karras hierarchy numbers them (the root 0, a left child
// by the last leaf of its range, a right child by the first), which gives the n - 1 internal
// nodes of any ordered binary tree the indices 0..n-2: an index k could only be claimed by
// the left child of the split after leaf k and by the right child of the split after leaf
// k - 1, and when the first holds both leaves k - 1 and k, the second lies inside it and is
// the single leaf k.
static void balanced_hierarchy(uint32_t n, std::vector<uint32_t>& child, std::vector<uint32_t>& lo,
                               std::vector<uint32_t>& hi, std::vector<uint32_t>& parent) {
  const uint32_t ni = n - 1;
  child.assign(2 * (size_t)ni, 0u);
  lo.assign(ni, 0u);
  hi.assign(ni, 0u);
  parent.assign((size_t)ni + n, 0u);
  struct Range {
    uint32_t idx, lo, hi;
  };
  std::vector<Range> todo{{0u, 0u, n - 1}};
  while (!todo.empty()) {
    const Range r = todo.back();
    todo.pop_back();
    const uint32_t g = r.lo + (r.hi - r.lo) / 2;
    lo[r.idx] = r.lo;
    hi[r.idx] = r.hi;
    uint32_t left, right;
    if (g == r.lo) {
      left = kLeafBit | g;
      parent[ni + g] = r.idx;
    } else {
      left = g;
      parent[g] = r.idx;
      todo.push_back({g, r.lo, g});
    }
    if (g + 1 == r.hi) {
      right = kLeafBit | (g + 1);
      parent[ni + g + 1] = r.idx;
    } else {
      right = g + 1;
      parent[g + 1] = r.idx;
      todo.push_back({g + 1, g + 1, r.hi});
    }
    child[2 * (size_t)r.idx] = left;
    child[2 * (size_t)r.idx + 1] = right;
  }
}

// The scene's BVH: the clustering build, or with `balanced` (the fallback when the clustering
// tree is too deep for the traversal stack) the balanced hierarchy over the Morton order,
// collapsed greedily.  Either tree gives the same hits; only traversal cost differs.
static int32_t update_scene_scale(rs_ctx* ctx, hipStream_t st);

static int32_t build_bvh(rs_ctx* ctx, hipStream_t st, bool balanced = false) {
  uint32_t n = ctx->numTris;
  const bool cluster = kUseClustering && !balanced;
  ctx->bvhBuild = balanced ? RS_BVH_BUILD_BALANCED : RS_BVH_BUILD_CLUSTERING;
  if (n == 0) {
    ctx->bvhRoot = 0;
    ctx->sceneScale = 0.f;
    return RS_OK;
  }
  // Scene box on the host (positions are uploaded from the host anyway).
  std::vector<float> pos(3 * (size_t)ctx->numVerts);
  HIP_TRY(hipMemcpy(pos.data(), ctx->positions.ptr, pos.size() * 4, hipMemcpyDeviceToHost));
  vec3 lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
  for (uint32_t i = 0; i < ctx->numVerts; i++) {
    vec3 p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  vec3 ext = hi - lo;
  vec3 inv{ext.x > 0.f ? 1.f / ext.x : 0.f, ext.y > 0.f ? 1.f / ext.y : 0.f, ext.z > 0.f ? 1.f / ext.z : 0.f};
  size_t ni = n > 1 ? n - 1 : 1;
  // scratch layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t oMorton = take(4 * (size_t)n), oMorton2 = take(4 * (size_t)n), oOrder = take(4 * (size_t)n),
         oOrder2 = take(4 * (size_t)n), oChild = take(8 * ni), oLo = take(4 * ni), oHi = take(4 * ni),
         oParent = take(4 * (ni + n)), oVisits = take(4 * ni), oBox = take(24 * ni);
  size_t oWorkA = take(4 * ni), oWorkB = take(4 * ni), oCnt = take(4 * (ni + 1)), oOffs = take(4 * (ni + 1)),
         oTmpRef = take(16 * ni), oTmpBox = take(96 * ni), oTmpSrc = take(16 * ni), oDp = take(16 * ni);
  // clustering build
  size_t oPBox = take(48 * (size_t)n), oPCount = take(8 * (size_t)n), oPChild = take(8 * (size_t)n),
         oPCl = take(4 * (size_t)n), oPClOut = take(4 * (size_t)n), oPNn = take(4 * (size_t)n),
         oPMerge = take(4 * ((size_t)n + 1)), oPMergeOffs = take(4 * ((size_t)n + 1)),
         oPKeep = take(4 * ((size_t)n + 1)), oPKeepOffs = take(4 * ((size_t)n + 1)), oPOffset = take(8 * (size_t)n),
         oPTris = take(sizeof(BvhTri) * (size_t)n);
  size_t sortTemp = 0, scanTemp = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, sortTemp, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (uint32_t*)nullptr, (int)n, 0, 30, st);
  hipcub::DeviceScan::ExclusiveSum(nullptr, scanTemp, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)(n + 1), st);
  size_t oTemp = take(std::max(sortTemp, scanTemp));
  if (ensure(ctx, ctx->bvhScratch, off) != RS_OK) return RS_E_OUT_OF_MEMORY;
  if (ensure(ctx, ctx->bvhNodes, sizeof(BvhNode) * ni) != RS_OK) return RS_E_OUT_OF_MEMORY;
  if (ensure(ctx, ctx->bvh4Nodes, sizeof(Bvh4Node) * ni) != RS_OK) return RS_E_OUT_OF_MEMORY;
  if (ensure(ctx, ctx->bvh4Src, 16 * ni) != RS_OK) return RS_E_OUT_OF_MEMORY;
  ctx->refitReady = false;
  // the walks address leaf triangles by 32-bit byte offsets (load_tri)
  if ((uint64_t)sizeof(BvhTri) * n > 0xffffffffull)
    return fail(ctx, RS_E_UNSUPPORTED, "scene of more triangles than 32-bit triangle offsets reach");
  if (ensure(ctx, ctx->bvhTris, sizeof(BvhTri) * (size_t)n) != RS_OK) return RS_E_OUT_OF_MEMORY;
  char* base = (char*)ctx->bvhScratch.ptr;
  LbvhBuild b;
  b.n = n;
  b.positions = (const float*)ctx->positions.ptr;
  b.indices = (const uint32_t*)ctx->indices.ptr;
  b.sceneMin = lo;
  b.sceneExtentInv = inv;
  b.morton = (uint32_t*)(base + oMorton);
  b.order = (uint32_t*)(base + oOrder);
  b.child = (uint32_t*)(base + oChild);
  b.rangeLo = (uint32_t*)(base + oLo);
  b.rangeHi = (uint32_t*)(base + oHi);
  b.parent = (uint32_t*)(base + oParent);
  b.visits = (uint32_t*)(base + oVisits);
  b.nodeBox = (float*)(base + oBox);
  b.nodes = (BvhNode*)ctx->bvhNodes.ptr;
  // Leaves in Morton order; the clustering build reorders them into ctx->bvhTris.
  b.tris = (n > kLeafMax && cluster) ? (BvhTri*)(base + oPTris) : (BvhTri*)ctx->bvhTris.ptr;
  uint32_t g = (n + 255) / 256;
  k_lbvh_morton<<<g, 256, 0, st>>>(b);
  HIP_TRY(hipcub::DeviceRadixSort::SortPairs(base + oTemp, sortTemp, b.morton, (uint32_t*)(base + oMorton2), b.order,
                                             (uint32_t*)(base + oOrder2), (int)n, 0, 30, st));
  b.morton = (uint32_t*)(base + oMorton2);
  b.order = (uint32_t*)(base + oOrder2);
  k_lbvh_leaves<<<g, 256, 0, st>>>(b);
  if (n == 1) {
    ctx->bvhRoot = kLeafBit | 0u;
  } else if (n <= kLeafMax) {
    ctx->bvhRoot = kLeafBit | (n - 1);
  } else {
    uint32_t gi = (uint32_t)((ni + 255) / 256);
    uint32_t rootBin = 0;
    if (!cluster) {
      if (balanced) {
        std::vector<uint32_t> hc, hl, hh, hp;
        balanced_hierarchy(n, hc, hl, hh, hp);
        HIP_TRY(hipMemcpy(b.child, hc.data(), 4 * hc.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b.rangeLo, hl.data(), 4 * hl.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b.rangeHi, hh.data(), 4 * hh.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(b.parent, hp.data(), 4 * hp.size(), hipMemcpyHostToDevice));
        k_fill_u32<<<gi, 256, 0, st>>>(b.visits, ni, 0u);
      } else {
        k_lbvh_hierarchy<<<gi, 256, 0, st>>>(b);
      }
      k_lbvh_refit<<<g, 256, 0, st>>>(b);
      k_lbvh_encode<<<gi, 256, 0, st>>>(b);
    } else {
      PlocBuild pb;
      pb.n = n;
      pb.box = (float*)(base + oPBox);
      pb.count = (uint32_t*)(base + oPCount);
      pb.child = (uint32_t*)(base + oPChild);
      pb.cl = (uint32_t*)(base + oPCl);
      pb.clOut = (uint32_t*)(base + oPClOut);
      pb.nn = (uint32_t*)(base + oPNn);
      pb.mergeFlag = (uint32_t*)(base + oPMerge);
      pb.mergeOffs = (uint32_t*)(base + oPMergeOffs);
      pb.keepFlag = (uint32_t*)(base + oPKeep);
      pb.keepOffs = (uint32_t*)(base + oPKeepOffs);
      pb.offset = (uint32_t*)(base + oPOffset);
      pb.trisIn = b.tris;
      pb.trisOut = (BvhTri*)ctx->bvhTris.ptr;
      pb.nodes = b.nodes;
      pb.m = n;
      pb.nodeBase = n;
      k_ploc_init<<<g, 256, 0, st>>>(pb);
      std::vector<std::pair<uint32_t, uint32_t>> created;  // (first id, count) per iteration
      uint32_t m = n, nodeBase = n;
      while (m > 1) {
        if (created.size() >= kMaxClusterIterations) return fail(ctx, RS_E_HIP, "BVH clustering did not converge");
        pb.m = m;
        pb.nodeBase = nodeBase;
        uint32_t gm = (m + 255) / 256, gm1 = (m + 1 + 255) / 256;
        k_ploc_nearest<<<gm, 256, 0, st>>>(pb);
        k_ploc_flags<<<gm1, 256, 0, st>>>(pb);
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(base + oTemp, scanTemp, pb.mergeFlag, pb.mergeOffs, (int)(m + 1), st));
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(base + oTemp, scanTemp, pb.keepFlag, pb.keepOffs, (int)(m + 1), st));
        k_ploc_merge<<<gm, 256, 0, st>>>(pb);
        uint32_t merged = 0, kept = 0;
        HIP_TRY(hipMemcpyAsync(&merged, pb.mergeOffs + m, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(&kept, pb.keepOffs + m, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (merged == 0 || kept != m - merged) return fail(ctx, RS_E_HIP, "BVH clustering made no progress");
        created.push_back({nodeBase, merged});
        nodeBase += merged;
        m = kept;
        std::swap(pb.cl, pb.clOut);
      }
      if (nodeBase != 2 * n - 1) return fail(ctx, RS_E_HIP, "BVH clustering produced a malformed tree");
      const uint32_t root = nodeBase - 1;
      static const uint32_t kZero = 0;
      HIP_TRY(hipMemcpyAsync(pb.offset + root, &kZero, 4, hipMemcpyHostToDevice, st));
      for (size_t it = created.size(); it-- > 0;)
        k_ploc_offsets<<<(created[it].second + 255) / 256, 256, 0, st>>>(pb, created[it].first, created[it].second);
      k_ploc_reorder<<<g, 256, 0, st>>>(pb);
      k_ploc_encode<<<gi, 256, 0, st>>>(pb);
#if !RS_BVH4_GREEDY
      for (const auto& it : created)
        k_bvh4_cost<<<(it.second + 255) / 256, 256, 0, st>>>(pb, (float*)(base + oDp), it.first, it.second);
#endif
      rootBin = root - n;
      ctx->buildIterations = (uint32_t)created.size();
      ctx->ploc = pb;
      ctx->plocIters = created;
      ctx->refitReady = true;
    }
    ctx->bvhRoot = 0;
    // Collapse to the 4-wide tree, one level per pass, starting from the binary root.
    Bvh4Collapse c;
    c.bin = b.nodes;
    c.out = (Bvh4Node*)ctx->bvh4Nodes.ptr;
    uint32_t* work[2] = {(uint32_t*)(base + oWorkA), (uint32_t*)(base + oWorkB)};
    c.cnt = (uint32_t*)(base + oCnt);
    c.offs = (const uint32_t*)(base + oOffs);
    c.tmpRef = (uint32_t*)(base + oTmpRef);
    c.tmpBox = (float*)(base + oTmpBox);
    c.tmpSrc = (uint32_t*)(base + oTmpSrc);
    c.src = (uint32_t*)ctx->bvh4Src.ptr;
    HIP_TRY(hipMemcpyAsync(work[0], &rootBin, 4, hipMemcpyHostToDevice, st));
    uint32_t count = 1, levelBase = 0, levels = 0;
    while (count > 0) {
      if (levelBase + count > ni) return fail(ctx, RS_E_HIP, "4-wide collapse overflowed its node array");
      c.workIn = work[levels & 1];
      c.workOut = work[(levels + 1) & 1];
      c.count = count;
      c.levelBase = levelBase;
      if (cluster && !RS_BVH4_GREEDY)
        k_bvh4_select_sah<<<(count + 1 + 255) / 256, 256, 0, st>>>(c, ctx->ploc, (const float*)(base + oDp));
      else
        k_bvh4_select<<<(count + 1 + 255) / 256, 256, 0, st>>>(c);
      HIP_TRY(hipcub::DeviceScan::ExclusiveSum(base + oTemp, scanTemp, c.cnt, (uint32_t*)(base + oOffs),
                                                (int)(count + 1), st));
      k_bvh4_emit<<<(count + 255) / 256, 256, 0, st>>>(c);
      uint32_t next = 0;
      HIP_TRY(hipMemcpyAsync(&next, (uint32_t*)(base + oOffs) + count, 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      levelBase += count;
      count = next;
      levels++;
    }
    ctx->bvh4Count = levelBase;
    ctx->bvh4Levels = levels;
    // the walks address nodes by 32-bit byte offsets (load_node4)
    if ((uint64_t)levelBase * sizeof(Bvh4Node) > 0xffffffffull)
      return fail(ctx, RS_E_UNSUPPORTED, "BVH of more nodes than 32-bit node offsets reach");
    if (levels > bvh_level_limit()) {
      if (!balanced) return build_bvh(ctx, st, true);
      return fail(ctx, RS_E_UNSUPPORTED, "BVH deeper than the traversal stack supports");
    }
  }
  HIP_TRY(hipGetLastError());
  return update_scene_scale(ctx, st);
}

// Light collection of the current host positions: every triangle whose material emits,
// in triangle order, and its power alias table (LightCollection / EmissivePowerSampler).
static int32_t build_light_collection(rs_ctx* ctx, const float* P) {
  const uint32_t numTris = ctx->numTris;
  std::vector<int32_t> triToLight(numTris, -1);
  std::vector<EmissiveTriangle> lights;
  std::vector<float> flux;
  for (uint32_t t = 0; t < numTris; t++) {
    const rs_material& m = ctx->hMaterials[ctx->hTriMaterial[t]];
    vec3 e{m.emissive[0], m.emissive[1], m.emissive[2]};
    if (!(e.x > 0.f || e.y > 0.f || e.z > 0.f)) continue;
    uint32_t i0 = ctx->hIndices[3 * t], i1 = ctx->hIndices[3 * t + 1], i2 = ctx->hIndices[3 * t + 2];
    EmissiveTriangle et{};
    et.p0 = vec3{P[3 * i0], P[3 * i0 + 1], P[3 * i0 + 2]};
    et.p1 = vec3{P[3 * i1], P[3 * i1 + 1], P[3 * i1 + 2]};
    et.p2 = vec3{P[3 * i2], P[3 * i2 + 1], P[3 * i2 + 2]};
    vec3 cr = cross(et.p1 - et.p0, et.p2 - et.p0);
    float len = length(cr);
    if (!(len > 0.f)) continue;
    et.area = 0.5f * len;
    et.normal = cr / len;
    et.prim = t;
    et.materialID = ctx->hTriMaterial[t];
    et.Le = e;
    // flux = luminance(average radiance) * area * pi (FinalizeIntegration.cs.slang:108).
    et.flux = luminance(e) * et.area * kPi;
    triToLight[t] = (int32_t)lights.size();
    lights.push_back(et);
    flux.push_back(et.flux);
  }
  ctx->numLights = (uint32_t)lights.size();
  std::vector<uvec2> table;
  double sum = 1.0;
  if (!lights.empty()) build_alias_table(flux, table, sum);
  ctx->invWeightsSum = lights.empty() ? 0.f : (float)(1.0 / sum);
  if (ensure(ctx, ctx->lights, sizeof(EmissiveTriangle) * std::max<size_t>(1, lights.size())) ||
      ensure(ctx, ctx->alias, sizeof(uvec2) * std::max<size_t>(1, table.size())))
    return RS_E_OUT_OF_MEMORY;
  HIP_TRY(hipMemcpy(ctx->triToLight.ptr, triToLight.data(), 4 * triToLight.size(), hipMemcpyHostToDevice));
  if (!lights.empty()) {
    HIP_TRY(hipMemcpy(ctx->lights.ptr, lights.data(), sizeof(EmissiveTriangle) * lights.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ctx->alias.ptr, table.data(), sizeof(uvec2) * table.size(), hipMemcpyHostToDevice));
  }
  return RS_OK;
}

// New boxes for new vertex positions over the kept topology (see k_refit_*).
// The scene's largest vertex coordinate magnitude, for the walks' culling slack
// (cull_slack); refreshed whenever the tree is built or refitted.
static int32_t update_scene_scale(rs_ctx* ctx, hipStream_t st) {
  ctx->sceneScale = 0.f;
  const size_t n = 3 * (size_t)ctx->numVerts;
  if (n == 0) return RS_OK;
  if (ensure(ctx, ctx->scaleBits, 4) != RS_OK) return RS_E_OUT_OF_MEMORY;
  uint32_t* bits = (uint32_t*)ctx->scaleBits.ptr;
  k_fill_u32<<<1, 64, 0, st>>>(bits, 1, 0u);
  const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 1024);
  k_max_abs_bits<<<blocks, 256, 0, st>>>((const float*)ctx->positions.ptr, n, bits);
  uint32_t h = 0;
  HIP_TRY(hipMemcpyAsync(&h, bits, 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  memcpy(&ctx->sceneScale, &h, 4);
  return RS_OK;
}

static int32_t refit_bvh(rs_ctx* ctx, hipStream_t st) {
  const uint32_t n = ctx->numTris;
  if (n == 0) return RS_OK;
  const uint32_t g = (n + 255) / 256;
  k_refit_tris<<<g, 256, 0, st>>>((BvhTri*)ctx->bvhTris.ptr, n, (const float*)ctx->positions.ptr,
                                  (const uint32_t*)ctx->indices.ptr);
  if (n > kLeafMax) {
    const PlocBuild& pb = ctx->ploc;
    k_refit_leaf_boxes<<<g, 256, 0, st>>>(pb);
    for (const auto& it : ctx->plocIters)
      k_refit_nodes<<<(it.second + 255) / 256, 256, 0, st>>>(pb, it.first, it.second);
    k_refit_bvh4<<<(ctx->bvh4Count + 255) / 256, 256, 0, st>>>((Bvh4Node*)ctx->bvh4Nodes.ptr,
                                                               (const uint32_t*)ctx->bvh4Src.ptr, pb.box,
                                                               ctx->bvh4Count);
  }
  HIP_TRY(hipGetLastError());
  return update_scene_scale(ctx, st);
}

// Drops the frozen replay snapshot: it holds material ids and hits of the old geometry.
static void drop_frozen_snapshot(rs_ctx* ctx) {
  ctx->frozenW = ctx->frozenH = 0;
  ctx->frozenMaterials = 0;
}

int32_t rs_set_scene(rs_ctx* ctx, const rs_scene_desc* sd) {
  if (!ctx || !sd) return RS_E_INVALID_ARG;
  if (!sd->positions || !sd->normals || !sd->indices || !sd->triMaterial || !sd->materials)
    return fail(ctx, RS_E_INVALID_ARG, "scene arrays must be non-null");
  if (sd->numMaterials == 0) return fail(ctx, RS_E_INVALID_ARG, "scene needs at least one material");
  // vertex and index arrays are addressed by 32-bit byte offsets (load3, load_indices)
  if ((uint64_t)12 * sd->numVertices > 0xffffffffull || (uint64_t)12 * sd->numTriangles > 0xffffffffull)
    return fail(ctx, RS_E_UNSUPPORTED, "scenes of more than 357,913,941 vertices or triangles are not supported");
  for (uint32_t i = 0; i < 3 * sd->numTriangles; i++)
    if (sd->indices[i] >= sd->numVertices) return fail(ctx, RS_E_INVALID_ARG, "vertex index out of range");
  for (uint32_t i = 0; i < sd->numTriangles; i++)
    if (sd->triMaterial[i] >= sd->numMaterials) return fail(ctx, RS_E_INVALID_ARG, "material index out of range");
  // Instances: disjoint vertex ranges, triangle ranges inside the soup whose triangles index
  // only their own instance's vertices.
  std::vector<rs_instance_desc> inst;
  if (sd->instances && sd->numInstances) inst.assign(sd->instances, sd->instances + sd->numInstances);
  else inst.push_back(rs_instance_desc{0, sd->numVertices, 0, sd->numTriangles});
  if (inst.size() > (1u << 29)) return fail(ctx, RS_E_INVALID_ARG, "too many instances");
  std::vector<uint32_t> vertInst(sd->numVertices, kNoInstance);
  for (uint32_t k = 0; k < (uint32_t)inst.size(); k++) {
    const rs_instance_desc& d = inst[k];
    if ((uint64_t)d.firstVertex + d.numVertices > sd->numVertices ||
        (uint64_t)d.firstTriangle + d.numTriangles > sd->numTriangles)
      return fail(ctx, RS_E_INVALID_ARG, "instance range outside the scene");
    for (uint32_t v = d.firstVertex; v < d.firstVertex + d.numVertices; v++) {
      if (vertInst[v] != kNoInstance) return fail(ctx, RS_E_INVALID_ARG, "instance vertex ranges overlap");
      vertInst[v] = k;
    }
    for (uint32_t t = d.firstTriangle; t < d.firstTriangle + d.numTriangles; t++)
      for (int j = 0; j < 3; j++)
        if (vertInst[sd->indices[3 * t + j]] != k)
          return fail(ctx, RS_E_INVALID_ARG, "an instance triangle indexes a vertex outside its instance");
  }
  if (int32_t e = ensure_device(ctx)) return e;
  // From here on the old scene is gone: until the build below succeeds the context has no
  // scene, so a failure cannot leave half-built buffers reachable by a frame.
  ctx->hasScene = false;
  ctx->refitReady = false;
  drop_frozen_snapshot(ctx);
  ctx->numTris = sd->numTriangles;
  ctx->numVerts = sd->numVertices;
  ctx->numMaterials = sd->numMaterials;
  ctx->diffuseModel = sd->diffuseModel;
  size_t vb = 12 * (size_t)sd->numVertices, tb = 12 * (size_t)sd->numTriangles;
  if (ensure(ctx, ctx->positions, vb) || ensure(ctx, ctx->normals, vb) || ensure(ctx, ctx->indices, tb) ||
      ensure(ctx, ctx->triMaterial, 4 * (size_t)sd->numTriangles) ||
      ensure(ctx, ctx->triToLight, 4 * (size_t)std::max(1u, sd->numTriangles)) ||
      ensure(ctx, ctx->materials, sizeof(Material) * sd->numMaterials) ||
      ensure(ctx, ctx->objPositions, vb) || ensure(ctx, ctx->objNormals, vb) ||
      ensure(ctx, ctx->vertInstance, 4 * (size_t)std::max(1u, sd->numVertices)) ||
      ensure(ctx, ctx->instTri, 8 * inst.size()) || ensure(ctx, ctx->instXforms, 48 * inst.size()))
    return RS_E_OUT_OF_MEMORY;
  HIP_TRY(hipMemcpy(ctx->positions.ptr, sd->positions, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->normals.ptr, sd->normals, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->objPositions.ptr, sd->positions, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->objNormals.ptr, sd->normals, vb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->indices.ptr, sd->indices, tb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->triMaterial.ptr, sd->triMaterial, 4 * (size_t)sd->numTriangles, hipMemcpyHostToDevice));
  static_assert(sizeof(Material) == sizeof(rs_material), "material layout");
  HIP_TRY(hipMemcpy(ctx->materials.ptr, sd->materials, sizeof(Material) * sd->numMaterials, hipMemcpyHostToDevice));
  if (sd->numVertices)
    HIP_TRY(hipMemcpy(ctx->vertInstance.ptr, vertInst.data(), 4 * (size_t)sd->numVertices, hipMemcpyHostToDevice));
  std::vector<uvec2> instTri(inst.size());
  for (size_t k = 0; k < inst.size(); k++) instTri[k] = uvec2{inst[k].firstTriangle, inst[k].numTriangles};
  HIP_TRY(hipMemcpy(ctx->instTri.ptr, instTri.data(), 8 * inst.size(), hipMemcpyHostToDevice));
  ctx->instances = inst;
  ctx->hIndices.assign(sd->indices, sd->indices + 3 * (size_t)sd->numTriangles);

  ctx->hTriMaterial.assign(sd->triMaterial, sd->triMaterial + sd->numTriangles);
  ctx->hMaterials.assign(sd->materials, sd->materials + sd->numMaterials);
  if (int32_t e = build_light_collection(ctx, sd->positions)) return e;
  int32_t r = build_bvh(ctx, nullptr);
  if (r != RS_OK) return r;
  ctx->hasScene = true;
  ctx->resetPending = true;
  return RS_OK;
}

// Geometry updates keep the scene only when they succeed: a failed update leaves positions,
// lights or tree out of step with each other, so the context drops the scene.
static int32_t geometry_update_result(rs_ctx* ctx, int32_t r) {
  if (r != RS_OK) {
    ctx->hasScene = false;
    ctx->refitReady = false;
  }
  return r;
}

int32_t rs_update_vertices(rs_ctx* ctx, const float* positions, const float* normals, uint32_t numVertices) {
  if (!ctx || !positions || !normals) return RS_E_INVALID_ARG;
  if (!ctx->hasScene) return fail(ctx, RS_E_NO_SCENE, "no scene");
  if (numVertices != ctx->numVerts) return fail(ctx, RS_E_SIZE_MISMATCH, "vertex count differs from the scene");
  drop_frozen_snapshot(ctx);
  size_t vb = 12 * (size_t)numVertices;
  int32_t r = RS_OK;
  hipError_t e = hipMemcpy(ctx->positions.ptr, positions, vb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ctx->normals.ptr, normals, vb, hipMemcpyHostToDevice);
  if (e != hipSuccess) r = fail(ctx, RS_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  if (r == RS_OK) r = build_light_collection(ctx, positions);
  if (r == RS_OK) r = (ctx->refitReady || ctx->numTris <= kLeafMax) ? refit_bvh(ctx, nullptr) : build_bvh(ctx, nullptr);
  return geometry_update_result(ctx, r);
}

int32_t rs_rebuild_bvh(rs_ctx* ctx) {
  if (!ctx) return RS_E_INVALID_ARG;
  if (!ctx->hasScene) return fail(ctx, RS_E_NO_SCENE, "no scene");
  return geometry_update_result(ctx, build_bvh(ctx, nullptr));
}

// New instance transforms: vertices, then the emissive triangles and their powers, then the
// alias table (on the host, from the powers) and the BVH refit.
static int32_t update_instances(rs_ctx* ctx, const float* xforms, uint32_t n) {
  const hipStream_t st = nullptr;
  HIP_TRY(hipMemcpy(ctx->instXforms.ptr, xforms, 48 * (size_t)n, hipMemcpyHostToDevice));
  const uint32_t nv = ctx->numVerts;
  if (nv)
    k_instance_transform<<<(nv + 255) / 256, 256, 0, st>>>(
        (const float*)ctx->objPositions.ptr, (const float*)ctx->objNormals.ptr, (const uint32_t*)ctx->vertInstance.ptr,
        (const float*)ctx->instXforms.ptr, nv, (float*)ctx->positions.ptr, (float*)ctx->normals.ptr);
  const uint32_t nl = ctx->numLights;
  if (nl) {
    if (ensure(ctx, ctx->lightFlux, 4 * (size_t)nl)) return RS_E_OUT_OF_MEMORY;
    k_update_lights<<<(nl + 255) / 256, 256, 0, st>>>((EmissiveTriangle*)ctx->lights.ptr, nl,
                                                      (const float*)ctx->positions.ptr,
                                                      (const uint32_t*)ctx->indices.ptr, (float*)ctx->lightFlux.ptr);
    HIP_TRY(hipGetLastError());
    std::vector<float> flux(nl);
    HIP_TRY(hipMemcpy(flux.data(), ctx->lightFlux.ptr, 4 * (size_t)nl, hipMemcpyDeviceToHost));
    bool degenerate = false;
    for (float f : flux) degenerate |= !(f >= 0.f);
    if (degenerate) {
      // An emissive triangle collapsed: the collection itself changes, rebuild it on the host.
      std::vector<float> pos(3 * (size_t)nv);
      HIP_TRY(hipMemcpy(pos.data(), ctx->positions.ptr, pos.size() * 4, hipMemcpyDeviceToHost));
      if (int32_t e = build_light_collection(ctx, pos.data())) return e;
    } else {
      std::vector<uvec2> table;
      double sum = 1.0;
      build_alias_table(flux, table, sum);
      ctx->invWeightsSum = (float)(1.0 / sum);
      HIP_TRY(hipMemcpy(ctx->alias.ptr, table.data(), sizeof(uvec2) * table.size(), hipMemcpyHostToDevice));
    }
  }
  HIP_TRY(hipGetLastError());
  if (ctx->refitReady || ctx->numTris <= kLeafMax) return refit_bvh(ctx, st);
  return build_bvh(ctx, st);
}

int32_t rs_update_instances(rs_ctx* ctx, const float* xforms, uint32_t n) {
  if (!ctx || !xforms) return RS_E_INVALID_ARG;
  if (!ctx->hasScene) return fail(ctx, RS_E_NO_SCENE, "no scene");
  if (n != (uint32_t)ctx->instances.size())
    return fail(ctx, RS_E_SIZE_MISMATCH, "one 3x4 transform per scene instance expected");
  for (size_t i = 0; i < 12 * (size_t)n; i++)
    if (!std::isfinite(xforms[i])) return fail(ctx, RS_E_INVALID_ARG, "instance transforms must be finite");
  drop_frozen_snapshot(ctx);
  return geometry_update_result(ctx, update_instances(ctx, xforms, n));
}

int32_t rs_set_env_map(rs_ctx* ctx, const rs_env_map_desc* d) {
  if (!ctx || !d || !d->texels || d->width == 0 || d->height == 0) return RS_E_INVALID_ARG;
  if (int32_t e = ensure_device(ctx)) return e;
  const size_t texBytes = 16 * (size_t)d->width * d->height;
  const size_t impTexels = env_mip_offset(kEnvImportanceLog2 + 1);
  if (ensure(ctx, ctx->envTexels, texBytes) || ensure(ctx, ctx->envImportance, 4 * impTexels))
    return RS_E_OUT_OF_MEMORY;
  HIP_TRY(hipMemcpy(ctx->envTexels.ptr, d->texels, texBytes, hipMemcpyHostToDevice));
  EnvMapView e{};
  e.texels = (const float*)ctx->envTexels.ptr;
  e.W = d->width;
  e.H = d->height;
  e.scale = vec3{d->intensity * d->tint[0], d->intensity * d->tint[1], d->intensity * d->tint[2]};
  env_rotation(d->rotation, e.toWorld, e.toLocal);
  float* imp = (float*)ctx->envImportance.ptr;
  e.importance = imp;
  k_env_importance<<<dim3(kEnvImportanceDim / 256, kEnvImportanceDim), 256>>>(e, imp);
  for (uint32_t k = 1; k <= kEnvImportanceLog2; k++) {
    const uint32_t dim = kEnvImportanceDim >> k;
    k_env_mip<<<(dim * dim + 255) / 256, 256>>>(imp + env_mip_offset(k - 1), imp + env_mip_offset(k), dim * 2);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  e.present = 1;
  ctx->env = e;
  return RS_OK;
}

int32_t rs_clear_env_map(rs_ctx* ctx) {
  if (!ctx) return RS_E_INVALID_ARG;
  ctx->env = EnvMapView{};
  return RS_OK;
}

int32_t rs_reset(rs_ctx* ctx) {
  if (!ctx) return RS_E_INVALID_ARG;
  if (ctx->sceneFrozen) return RS_OK;  // reset() leaves a frozen scene's frame alone (:813-820)
  ctx->frameCount = 0;
  ctx->seed = 0;
  ctx->resetPending = true;
  return RS_OK;
}

int32_t rs_set_timing(rs_ctx* ctx, int enable) {
  if (!ctx) return RS_E_INVALID_ARG;
  ctx->timing = enable != 0;
  if (ctx->timing && !ctx->ev[0])
    for (auto& e : ctx->ev) HIP_TRY(hipEventCreate(&e));
  return RS_OK;
}

int32_t rs_get_stats(rs_ctx* ctx, rs_stats* out) {
  if (!ctx || !out) return RS_E_INVALID_ARG;
  unsigned long long r[2] = {0, 0};
  if (ctx->rays.ptr) HIP_TRY(hipMemcpy(r, ctx->rays.ptr, sizeof(r), hipMemcpyDeviceToHost));
  ctx->stats.raysClosestHit = r[0];
  ctx->stats.raysVisibility = r[1];
  ctx->stats.frameCount = ctx->frameCount;
  ctx->stats.seed = ctx->seed;
  ctx->stats.bvhNodes = ctx->bvh4Count;
  ctx->stats.bvhLevels = ctx->bvh4Levels;
  ctx->stats.bvhBuild = ctx->bvhBuild;
  if (ctx->timing) {
    HIP_TRY(hipEventSynchronize(ctx->ev[9]));
    float ms = 0.f;
    for (int i = 0; i < 8; i++) {
      hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]);
      ctx->stats.msStage[i] = ms;
    }
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[9]);
    ctx->stats.msTotal = ms;
  }
  *out = ctx->stats;
  return RS_OK;
}

// What one frame needs, from the options, the camera and the frame size.
struct FramePlan {
  uint32_t W = 0, H = 0;
  size_t P = 0, KPP = 0;
  uint32_t mode = 0, KP = 1;
  uint32_t stride = 0;  // tile-order positions of a per-pixel launch (blocks x 256)
  dim3 grid;
  bool collects = false, robust = false, tasks = false, backup = false, spatial = false;
  uint32_t neighbors = 0;
};

static int32_t plan_frame(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, FramePlan& f) {
  if (io->width == 0 || io->height == 0) return fail(ctx, RS_E_SIZE_MISMATCH, "zero frame size");
  if (io->width > 4096 || io->height > 4096)
    return fail(ctx, RS_E_FRAME_TOO_LARGE, "Frame dimensions up to 4096 pixels width/height are supported.");
  if (io->d_vbuffer && io->vbufferPrimitiveIndexBits > 28)
    return fail(ctx, RS_E_INVALID_ARG, "vbufferPrimitiveIndexBits must be 0 (128-bit header) or 1..28");
  const rs_options& o = ctx->opt;
  // Refused before anything of the context changes, so that a failed frame (or batch frame)
  // leaves every context as it was.
  const uint32_t K = 1 + (uint32_t)std::ceil(cam->shutterSpeed / o.artificialFrameTime) + 1;
  if (K > (uint32_t)kMaxCameraKeyframes) return fail(ctx, RS_E_BAD_OPTION, "shutter / frame time needs too many keyframes");
  f.W = io->width;
  f.H = io->height;
  f.P = (size_t)f.W * f.H;
  // MultiScatter with a zero shutter has one partition's worth of time and runs as
  // ScatterOnly (every partition would splat to the same cell).
  f.mode = o.temporalReuse;
  if (f.mode == RS_TEMPORAL_MULTI_SCATTER && !(cam->shutterSpeed > 0.f)) f.mode = RS_TEMPORAL_SCATTER_ONLY;
  f.KP = f.mode == RS_TEMPORAL_MULTI_SCATTER ? o.numTimePartitions : 1u;
  f.KPP = (size_t)f.KP * f.P;
  const uint32_t tw = 1u << ctx->tileShift, th = 256u >> ctx->tileShift;
  f.grid = dim3((f.W + tw - 1) / tw, (f.H + th - 1) / th);
  f.stride = f.grid.x * f.grid.y * 256u;
  const bool temporal = o.enableTemporalResampling != 0;
  f.collects = temporal && (f.mode == RS_TEMPORAL_GATHER_ONLY || f.mode == RS_TEMPORAL_SCATTER_BACKUP);
  f.robust = f.collects && o.gatherMechanism == RS_GATHER_ROBUST;
  f.tasks = temporal && f.mode != RS_TEMPORAL_GATHER_ONLY;
  f.backup = f.mode == RS_TEMPORAL_SCATTER_BACKUP;
  f.spatial = o.enableSpatialResampling != 0;
  f.neighbors = o.neighborCount;
  return RS_OK;
}

// The context's own per-pixel buffers (history, frame outputs kept for inspection): sized
// for the frame, allocated when they grow unless the context may not allocate.
static int32_t prepare_state(rs_ctx* ctx, const FramePlan& f) {
  const size_t P = f.P, KPP = f.KPP;
  for (int i = 0; i < 5; i++) {
    if (i == 3 && !ctx->opt.captureStages) continue;
    if (i == 4 && !ctx->sceneFrozen) continue;
    if (ensure(ctx, ctx->R[i], 32 * P) || ensure(ctx, ctx->C[i], 112 * P)) return ctx->noAlloc ? RS_E_SIZE_MISMATCH : RS_E_OUT_OF_MEMORY;
  }
  if (ensure(ctx, ctx->mvec, 8 * P) || ensure(ctx, ctx->vbuffer, 12 * P) || ensure(ctx, ctx->floatingCoords, 8 * P) ||
      ensure(ctx, ctx->splatDst, 4 * KPP) || ensure(ctx, ctx->cellCount, 4 * KPP) || ensure(ctx, ctx->cellOffset, 4 * KPP) ||
      ensure(ctx, ctx->sorted, 4 * KPP) || ensure(ctx, ctx->color, 16 * P))
    return ctx->noAlloc ? RS_E_SIZE_MISMATCH : RS_E_OUT_OF_MEMORY;
  if (f.W != ctx->W || f.H != ctx->H) {
    ctx->W = f.W;
    ctx->H = f.H;
    ctx->resetPending = true;  // history of another size cannot be reused
  }
  return RS_OK;
}

// Scratch memory of one frame: device memory that only this frame's launches use.
struct QueueMem {
  uint32_t *flags = nullptr, *offsets = nullptr, *queue = nullptr, *fetch = nullptr;
  uint32_t* chunkSums = nullptr;  // the scan's per-chunk sums
};
struct FrameScratch {
  PackedShifted* shifted = nullptr;  // Robust gather: 8 shifted paths per pixel
  PackedCand* cand = nullptr;        // temporal candidate slots
  uint32_t* sortedCell = nullptr;    // cell of each splat-list entry
  uint32_t* splatRank = nullptr;     // rank of each splat entry in its cell; then the sort's temporary
  uint32_t* largeCells = nullptr;    // [0] count, then the cells too long for a thread to sort
  uint32_t* splatFetch = nullptr;    // task counter of the splat entries' own launch
  void* cellScanTemp = nullptr;
  size_t cellScanTempBytes = 0;
  QueueMem tq, sq, rq, mq;           // temporal, spatial, Robust-shift and MultiReproject task lists
  float* spatialM1 = nullptr;
  PackedCand* spatialCand = nullptr;
};

// One piece of the scratch: its size, where its pointer goes and the context buffer that
// holds it when the context owns its scratch.
struct ScratchPiece {
  size_t bytes;
  void** out;
  DevBuf* own;
};

static size_t scan_temp_bytes(size_t n) {
  size_t tb = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)std::max<size_t>(1, n));
  return tb;
}

static void queue_pieces(std::vector<ScratchPiece>& v, QueueBufs& qb, QueueMem& q, size_t n) {
  n = std::max<size_t>(1, n);
  v.push_back({4 * n, (void**)&q.flags, &qb.flags});
  v.push_back({4 * n, (void**)&q.offsets, &qb.offsets});
  v.push_back({4 * n, (void**)&q.queue, &qb.queue});
  v.push_back({4, (void**)&q.fetch, &qb.fetch});
  v.push_back({4 * ((n + kBatchScanChunk - 1) / kBatchScanChunk), (void**)&q.chunkSums, &qb.chunkSums});
}

static int32_t scratch_pieces(rs_ctx* ctx, const FramePlan& f, FrameScratch& s, std::vector<ScratchPiece>& v) {
  v.clear();
  const size_t P = f.P, KPP = f.KPP;
  if (f.robust) {
    v.push_back({sizeof(PackedShifted) * 8 * P, (void**)&s.shifted, &ctx->shifted});
    queue_pieces(v, ctx->robustQ, s.rq, f.stride);
  }
  if (f.tasks) {
    const size_t lists = f.backup ? 2 : 1;
    if (lists * f.stride > (size_t)INT32_MAX || KPP > (size_t)INT32_MAX)
      return fail(ctx, RS_E_FRAME_TOO_LARGE, "task list longer than 2^31 entries");
    v.push_back({sizeof(PackedCand) * ((f.backup ? 2 * P : P) + KPP), (void**)&s.cand, &ctx->cand});
    v.push_back({4 * KPP, (void**)&s.sortedCell, &ctx->sortedCell});
    v.push_back({4 * KPP, (void**)&s.splatRank, &ctx->splatRank});
    v.push_back({4 * (KPP / (kSortSmall + 1) + 1), (void**)&s.largeCells, &ctx->largeCells});
    v.push_back({4, (void**)&s.splatFetch, &ctx->splatFetch});
    s.cellScanTempBytes = scan_temp_bytes(KPP);
    v.push_back({s.cellScanTempBytes, &s.cellScanTemp, &ctx->scanTemp});
    queue_pieces(v, ctx->temporalQ, s.tq, lists * f.stride);
    if (f.mode == RS_TEMPORAL_MULTI_SCATTER) queue_pieces(v, ctx->multiQ, s.mq, f.stride);
  }
  if (f.spatial) {
    const size_t nP = (size_t)f.neighbors * P;
    if (2 * (size_t)f.neighbors * f.stride > (size_t)INT32_MAX)
      return fail(ctx, RS_E_FRAME_TOO_LARGE, "task list longer than 2^31 entries");
    v.push_back({4 * std::max<size_t>(1, nP), (void**)&s.spatialM1, &ctx->spatialM1});
    v.push_back({sizeof(PackedCand) * std::max<size_t>(1, nP), (void**)&s.spatialCand, &ctx->spatialCand});
    queue_pieces(v, ctx->spatialQ, s.sq, 2 * (size_t)f.neighbors * f.stride);
  }
  return RS_OK;
}

// The context's own scratch buffers, grown as needed.
static int32_t bind_owned_scratch(rs_ctx* ctx, std::vector<ScratchPiece>& v) {
  for (ScratchPiece& q : v) {
    if (ensure(ctx, *q.own, q.bytes)) return ctx->noAlloc ? RS_E_SIZE_MISMATCH : RS_E_OUT_OF_MEMORY;
    *q.out = q.own->ptr;
  }
  return RS_OK;
}

// Everything a frame allocates, before its first launch.
static int32_t prepare_frame(rs_ctx* ctx, const FramePlan& f, FrameScratch& s) {
  if (int32_t e = prepare_state(ctx, f)) return e;
  std::vector<ScratchPiece> pieces;
  if (int32_t e = scratch_pieces(ctx, f, s, pieces)) return e;
  return bind_owned_scratch(ctx, pieces);
}

static void fill_scene_view(const rs_ctx* ctx, SceneView& s) {
  s.positions = (const float*)ctx->positions.ptr;
  s.normals = (const float*)ctx->normals.ptr;
  s.indices = (const uint32_t*)ctx->indices.ptr;
  s.triMaterial = (const uint32_t*)ctx->triMaterial.ptr;
  s.triToLight = (const int32_t*)ctx->triToLight.ptr;
  s.materials = (const Material*)ctx->materials.ptr;
  s.lights = (const EmissiveTriangle*)ctx->lights.ptr;
  s.aliasTable = (const uvec2*)ctx->alias.ptr;
  s.numTriangles = ctx->numTris;
  s.numVertices = ctx->numVerts;
  s.numMaterials = ctx->numMaterials;
  s.numLights = ctx->numLights;
  s.invWeightsSum = ctx->invWeightsSum;
  s.diffuseModel = ctx->diffuseModel;
  s.env = ctx->env;
}

// A task list over scratch memory.
static TaskQueue task_queue(const QueueMem& m, uint32_t lists, uint32_t stride) {
  TaskQueue q;
  q.flags = m.flags;
  q.offsets = m.offsets;
  q.queue = m.queue;
  q.fetch = m.fetch;
  q.stride = stride;
  q.lists = lists;
  return q;
}

// After the flags kernel: scan, scatter the live entries, reset the fetch position.
static int32_t queue_build(rs_ctx* ctx, const QueueMem& m, const TaskQueue& q, const FrameParams& p, dim3 grid,
                           hipStream_t st) {
  const uint32_t n = q.lists * q.stride, chunks = (n + kBatchScanChunk - 1) / kBatchScanChunk;
  k_scan_reduce<<<chunks, 256, 0, st>>>(q.flags, n, m.chunkSums);
  k_scan_chunks<<<1, 256, 0, st>>>(m.chunkSums, chunks);
  k_scan_queue<<<chunks, 256, 0, st>>>(p, q, m.chunkSums, grid.x);
  return RS_OK;
}

// Grid of a persistent task kernel: as many 256-lane blocks as fit on the device at once.
static int32_t task_blocks(rs_ctx* ctx, const void* kernel, uint32_t& blocks) {
  if (blocks) return RS_OK;
  int perCU = 0, cus = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, 256, 0));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  blocks = (uint32_t)std::max(1, perCU * cus);
  return RS_OK;
}

// What the launches of one frame need, decided on the host: the kernel arguments of every
// phase (the temporal phase, then one set per spatial iteration, whose reservoir slots
// ping-pong), the stages that run and the copies in and out.
struct FrameRecord {
  KernelArgs a;                      // vbuffer .. temporal resampling
  std::vector<KernelArgs> spatial;   // spatial iteration i
  const PackedReservoir* finalR = nullptr;
  bool env = false, stats = false;
  bool overlap = false;              // stages run on the side stream beside the initial candidates
  bool temporal = false;             // temporal reuse runs (not the first frame, no reset)
  bool robust = false, collects = false, splat = false, multi = false, tasks = false, gather = false;
  uint32_t kind = 0;                 // TemporalTaskKind of the task kernels
  TaskLists tl{};
  SpatialTasks sp{};
  TaskQueue rq{};
  TaskQueue mq{};                    // MultiReproject's live previous samples
  const float* mvecIn = nullptr;     // caller's motion vectors, copied over the computed ones
  float* colorOut = nullptr;         // caller's colour output
  bool capture = false;              // copy the initial candidates to slot 3
  PackedReservoir* captureR = nullptr;
  PackedRecon* captureC = nullptr;
  uint32_t* rayCount = nullptr;
  uint32_t* pathLength = nullptr;
};

// beginFrame, the arguments of every launch and endFrame (execute, RS/ReservoirSplatting.cpp:
// 399-509): the frame's host side, before anything is launched.  The context's state moves
// on to the next frame here (seed, keyframe ring, history slots, frameCount).
static int32_t record_frame(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, const FramePlan& f,
                            const FrameScratch& s, bool overlap, FrameRecord& r) {
  r.overlap = overlap;
  const rs_options& o = ctx->opt;
  const uint32_t W = f.W, H = f.H, mode = f.mode, KP = f.KP;
  const bool frozen = ctx->sceneFrozen;

  // beginFrame: seed and keyframe count.
  ctx->seed = o.useFixedSeed ? o.fixedSeed : ctx->seed + 1;
  const uint32_t K = 1 + (uint32_t)std::ceil(cam->shutterSpeed / o.artificialFrameTime) + 1;  // <= kMaxCameraKeyframes (plan_frame)
  bool shouldReset = ctx->resetPending || (io->sceneFlags & RS_SCENE_SHOULD_RESET) != 0;
  if (K != ctx->prevNumKeyframes) shouldReset = true;
  ctx->prevNumKeyframes = K;
  ctx->cm.artificialFrameTime = o.artificialFrameTime;
  ctx->cm.numKeyframes = K;
  vec3 cpos{cam->position[0], cam->position[1], cam->position[2]};
  vec3 ctgt{cam->target[0], cam->target[1], cam->target[2]};
  if (ctx->frameCount == 0 || shouldReset) {
    for (uint32_t i = 0; i < K; i++) {
      ctx->cm.positions[i] = cpos;
      ctx->cm.targets[i] = ctgt;
    }
  } else {
    uint32_t ci = ctx->frameCount % K;
    ctx->cm.positions[ci] = cpos;
    ctx->cm.targets[ci] = ctgt;
  }
  ctx->resetPending = false;

  KernelArgs& a = r.a;
  std::memset(&a, 0, sizeof(a));
  r.env = ctx->env.present != 0;
  r.stats = io->d_rayCount != nullptr;
  FrameParams& p = a.p;
  p.W = W;
  p.H = H;
  p.frameCount = ctx->frameCount;
  p.seed = ctx->seed;
  p.useFixedSeed = o.useFixedSeed;
  p.fixedSeed = o.fixedSeed;
  p.spp = o.samplesPerPixel;
  p.maxSurfaceBounces = o.maxSurfaceBounces;
  p.maxDiffuseBounces = o.maxDiffuseBounces;
  p.maxSpecularBounces = o.maxSpecularBounces;
  p.maxTransmissionBounces = o.maxTransmissionBounces;
  p.useRussianRoulette = o.useRussianRoulette;
  p.misHeuristic = o.misHeuristic;
  p.misPowerExponent = o.misPowerExponent;
  p.specularRoughnessThreshold = o.specularRoughnessThreshold;
  p.adjustPrimary = o.gbufferAdjustShadingNormals;
  p.adjustSecondary = o.adjustShadingNormals;
  p.camera.posW = cpos;
  p.camera.up = vec3{0.f, 1.f, 0.f};
  p.camera.focalLength = cam->focalLength;
  p.camera.frameHeight = cam->frameHeight;
  p.camera.aspectRatio = cam->aspectRatio;
  p.camera.focalDistance = cam->focalDistance;
  p.camera.apertureRadius = cam->apertureRadius;
  p.camera.shutterSpeed = cam->shutterSpeed;
  p.cm = ctx->cm;
  p.quirkMask = o.quirkMask;
  p.tileShift = ctx->tileShift;  // the lane -> pixel map of every per-pixel kernel (f.grid's tiles)
  p.firstSample = (o.quirkMask & RS_QUIRK_SPP_LAST_SAMPLE) ? o.samplesPerPixel - 1 : 0u;
  fill_camera_cache(p);
  fill_scene_view(ctx, a.s);
  a.bvh.nodes = (const Bvh4Node*)ctx->bvh4Nodes.ptr;
  a.bvh.tris = (const BvhTri*)ctx->bvhTris.ptr;
  a.bvh.root = ctx->bvhRoot;
  a.bvh.numTris = ctx->numTris;
  a.bvh.scale = ctx->sceneScale;
  TemporalOptions& to = a.o;
  to.gatherMechanism = o.gatherMechanism;
  to.backupMIS = o.scatterBackupMIS;
  to.useConfidence = o.useConfidenceWeightsTemporally;
  to.spatialNeighbors = o.neighborCount;
  to.useConfidenceSpatial = o.useConfidenceWeightsSpatially;
  to.spatialRadius = o.spatialGatherRadius;
  to.numPartitions = KP;
  a.rays = (unsigned long long*)ctx->rays.ptr;

  PassBuffers& b = a.b;
  b.currR = (PackedReservoir*)ctx->R[ctx->curr].ptr;
  b.currC = (PackedRecon*)ctx->C[ctx->curr].ptr;
  b.prevR = (PackedReservoir*)ctx->R[ctx->prev].ptr;
  b.prevC = (PackedRecon*)ctx->C[ctx->prev].ptr;
  b.backupR = (PackedReservoir*)ctx->R[2].ptr;
  b.backupC = (PackedRecon*)ctx->C[2].ptr;
  b.mvec = (float*)ctx->mvec.ptr;
  b.vbuffer = (HitInfo*)ctx->vbuffer.ptr;
  b.floatingCoords = (float*)ctx->floatingCoords.ptr;
  b.splatDst = (uint32_t*)ctx->splatDst.ptr;
  b.cellCount = (uint32_t*)ctx->cellCount.ptr;
  b.cellOffset = (uint32_t*)ctx->cellOffset.ptr;
  b.splatRank = s.splatRank;
  b.sorted = (uint32_t*)ctx->sorted.ptr;
  b.neighborOffsets = (const int8_t*)ctx->neighborOffsets.ptr;
  b.color = (float*)ctx->color.ptr;
  b.vbufferIn = io->d_vbuffer;
  b.vbufferPrimBits = io->vbufferPrimitiveIndexBits;
  b.numInstances = (uint32_t)ctx->instances.size();
  b.instTri = (const uvec2*)ctx->instTri.ptr;
  b.rayCountOut = io->d_rayCount;

  r.mvecIn = (const float*)io->d_mvec;
  r.colorOut = (float*)io->d_color;
  r.rayCount = io->d_rayCount;
  r.pathLength = io->d_pathLength;
  r.capture = o.captureStages && ctx->R[3].ptr;
  r.captureR = (PackedReservoir*)ctx->R[3].ptr;
  r.captureC = (PackedRecon*)ctx->C[3].ptr;

  shouldReset |= (ctx->frameCount == 0);
  r.temporal = o.enableTemporalResampling && !shouldReset;
  if (r.temporal) {
    r.robust = f.robust;
    r.collects = f.collects;
    if (r.robust) {
      b.shiftedPaths = s.shifted;
      r.rq = task_queue(s.rq, 1, f.stride);
    }
    r.splat = mode != RS_TEMPORAL_GATHER_ONLY;
    r.multi = mode == RS_TEMPORAL_MULTI_SCATTER;
    if (r.multi) r.mq = task_queue(s.mq, 1, f.stride);
    r.gather = mode == RS_TEMPORAL_GATHER_ONLY;
    r.tasks = f.tasks;
    if (r.tasks) {
      r.tl.splatBase = (uint32_t)(f.backup ? 2 * f.P : f.P);
      r.tl.q = task_queue(s.tq, f.backup ? 2 : 1, f.stride);
      r.tl.cand = s.cand;
      r.tl.sortedCell = s.sortedCell;
      r.tl.splatFetch = s.splatFetch;
    }
    r.kind = mode == RS_TEMPORAL_SCATTER_BACKUP  ? kTasksBackup
             : mode == RS_TEMPORAL_MULTI_SCATTER ? kTasksMulti
                                                 : kTasksScatter;
  }
  // Spatial reuse ping-pongs with the history buffers, or, while the scene is frozen, with
  // the temporary buffers so that the history survives (RS/ReservoirSplatting.cpp:2003-2004).
  int& swapSlot = frozen ? ctx->tmp : ctx->prev;
  int temporalSlot = ctx->curr;
  r.spatial.clear();
  if (f.spatial) {
    r.sp = SpatialTasks{s.spatialM1, s.spatialCand, task_queue(s.sq, 2 * o.neighborCount, f.stride)};
    for (uint32_t it = 0; it < o.numSpatialIterations; it++) {
      std::swap(ctx->curr, swapSlot);
      temporalSlot = swapSlot;
      // spatial reads the swap slot (the temporal output) and writes `curr`.
      KernelArgs sa = a;
      sa.b.currR = (PackedReservoir*)ctx->R[ctx->curr].ptr;
      sa.b.currC = (PackedRecon*)ctx->C[ctx->curr].ptr;
      sa.b.prevR = (PackedReservoir*)ctx->R[swapSlot].ptr;
      sa.b.prevC = (PackedRecon*)ctx->C[swapSlot].ptr;
      sa.b.outR = sa.b.currR;
      sa.b.outC = sa.b.currC;
      sa.o.spatialIteration = it;
      r.spatial.push_back(sa);
    }
  }
  r.finalR = (const PackedReservoir*)ctx->R[ctx->curr].ptr;

  // The render-graph dictionary (:503-506).
  const bool replaying = (io->sceneFlags & RS_SCENE_REPLAYING) != 0;
  ctx->flags.suppressAutoReset = (frozen || replaying) ? 1u : 0u;
  ctx->flags.shouldAccumulate = (!replaying || (io->sceneFlags & RS_SCENE_ACCUMULATION_FRAME)) ? 1u : 0u;
  ctx->flags.viewAccumulation = (replaying && (io->sceneFlags & RS_SCENE_VIEW_ACCUMULATION)) ? 1u : 0u;
  // endFrame (:1648-1680): a frozen scene keeps frameCount and the history.
  ctx->splatPartitions = KP;
  ctx->splatValid = r.splat;
  ctx->lastParams = p;
  ctx->rendered = true;
  ctx->lastTemporal = temporalSlot;
  if (!frozen) {
    ctx->frameCount++;
    std::swap(ctx->curr, ctx->prev);
    ctx->lastFinal = ctx->prev;
  } else {
    ctx->lastFinal = ctx->curr;
  }
  ctx->temporalValid = o.enableSpatialResampling != 0;
  return RS_OK;
}

// Grids of the persistent task kernels for this context's device (once per context).
static int32_t frame_task_blocks(rs_ctx* ctx) {
  if (int32_t e = task_blocks(ctx, ctx->env.present ? (const void*)k_temporal_tasks<true> : (const void*)k_temporal_tasks<false>,
                              ctx->taskBlocks[ctx->env.present ? 1 : 0]))
    return e;
  if (int32_t e = task_blocks(ctx, ctx->env.present ? (const void*)k_spatial_tasks<true> : (const void*)k_spatial_tasks<false>,
                              ctx->spatialBlocks[ctx->env.present ? 1 : 0]))
    return e;
  if (int32_t e = task_blocks(ctx, ctx->env.present ? (const void*)k_multi_reproject_tasks<true>
                                                    : (const void*)k_multi_reproject_tasks<false>,
                              ctx->multiBlocks[ctx->env.present ? 1 : 0]))
    return e;
  return task_blocks(ctx, ctx->env.present ? (const void*)k_robust_tasks<true> : (const void*)k_robust_tasks<false>,
                     ctx->robustBlocks[ctx->env.present ? 1 : 0]);
}

// The launches of one recorded frame on one stream.
static int32_t launch_frame(rs_ctx* ctx, const FrameRecord& r, const FramePlan& f, const FrameScratch& s,
                            hipStream_t st) {
  const KernelArgs& a = r.a;
  const FrameParams& p = a.p;
  const PassBuffers& b = a.b;
  const bool env = r.env;
  const size_t P = f.P;
  const dim3 grid = f.grid, block(256);
  dim3 grid16((f.W + 15) / 16, (f.H + 15) / 16);
  auto mark = [&](int i) {
    if (ctx->timing) hipEventRecord(ctx->ev[i], st);
  };
  // Counters and splat cells are cleared by kernels rather than memsets, so a frame can be
  // recorded into a graph (rs_batch_render).
  k_fill_u32<<<1, 64, 0, st>>>((uint32_t*)ctx->rays.ptr, 8, 0u);

  mark(0);
  // Stages that read only the previous frame's samples and this frame's primary hits run on
  // the context's side stream while the initial candidates trace: the primary hits and motion
  // vectors, the Robust ring shifts, the collect, the reprojection, the splat sort and then the
  // splat-list entries' shifts.  `st` waits for the sort before the temporal flags and for the
  // splat shifts before the combine.  With stage timing on, every stage runs on `st` in order
  // so that the stage events bracket one stage each.  A frame recorded into a graph (`st`
  // capturing, as in the stream form of rs_batch_render) also keeps every stage on `st`:
  // recorded frames that forked onto the side stream ended the process inside
  // rs_batch_render (a host segmentation fault, no GPU fault), for a reason not found.
  const bool overlap = r.overlap;
  hipStream_t sd = st;
  // Joins the side stream back into `st` on every way out, an early error return included,
  // so that nothing the caller enqueues after the frame overlaps the frame's side work.
  struct SideJoin {
    hipEvent_t ev = nullptr;
    hipStream_t side = nullptr, main = nullptr;
    ~SideJoin() {
      if (side && hipEventRecord(ev, side) == hipSuccess) (void)hipStreamWaitEvent(main, ev, 0);
    }
  } sideJoin;
  if (overlap) {
    sd = ctx->side;
    HIP_TRY(hipEventRecord(ctx->forkEv, st));
    HIP_TRY(hipStreamWaitEvent(sd, ctx->forkEv, 0));
    sideJoin.ev = ctx->splatEv;
    sideJoin.side = sd;
    sideJoin.main = st;
  }
  // The primary-hit buffer is always produced (DoF circle of confusion): decoded from the
  // caller's VBuffer or traced; the caller's motion vectors, when given, replace the computed ones.
  auto vbuffer = [&](hipStream_t q) -> int32_t {
    k_vbuffer<<<grid, block, 0, q>>>(a);
    if (r.mvecIn) HIP_TRY(hipMemcpyAsync(b.mvec, r.mvecIn, 8 * P, hipMemcpyDeviceToDevice, q));
    return RS_OK;
  };
  if (!overlap)
    if (int32_t e = vbuffer(st)) return e;
  mark(1);
  if (r.stats) {
    if (env) k_initial<true, true><<<grid, block, 0, st>>>(a);
    else k_initial<false, true><<<grid, block, 0, st>>>(a);
  } else {
    if (env) k_initial<true, false><<<grid, block, 0, st>>>(a);
    else k_initial<false, false><<<grid, block, 0, st>>>(a);
  }
  if (r.capture) {
    HIP_TRY(hipMemcpyAsync(r.captureR, b.currR, 32 * P, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(r.captureC, b.currC, 112 * P, hipMemcpyDeviceToDevice, st));
  }
  mark(2);
  if (overlap)
    if (int32_t e = vbuffer(sd)) return e;
  if (r.temporal) {
    if (r.robust) {
      k_robust_flags<<<grid, block, 0, sd>>>(a, r.rq);
      if (int32_t e = queue_build(ctx, s.rq, r.rq, p, grid, sd)) return e;
      if (env) k_robust_tasks<true><<<ctx->robustBlocks[1], 256, 0, sd>>>(a, r.rq);
      else k_robust_tasks<false><<<ctx->robustBlocks[0], 256, 0, sd>>>(a, r.rq);
    }
    if (r.collects) k_collect<<<grid, block, 0, sd>>>(a);
    mark(3);
    const size_t KPP = f.KPP;
    const uint32_t linK = (uint32_t)((KPP + 255) / 256);
    if (r.splat) {
      k_fill_u32<<<linK, 256, 0, sd>>>(b.cellCount, KPP, 0u);
      if (!r.multi) {
        k_reproject<<<grid, block, 0, sd>>>(a);
      } else {
        k_multi_flags<<<grid, block, 0, sd>>>(a, r.mq, a.o.numPartitions);
        if (int32_t e = queue_build(ctx, s.mq, r.mq, p, grid, sd)) return e;
        if (env) k_multi_reproject_tasks<true><<<ctx->multiBlocks[1], 256, 0, sd>>>(a, r.mq, a.o.numPartitions);
        else k_multi_reproject_tasks<false><<<ctx->multiBlocks[0], 256, 0, sd>>>(a, r.mq, a.o.numPartitions);
      }
    }
    mark(4);
    if (r.tasks) {
      size_t tb = s.cellScanTempBytes;
      HIP_TRY(hipcub::DeviceScan::ExclusiveSum(s.cellScanTemp, tb, b.cellCount, b.cellOffset, (int)KPP, sd));
      k_splat_fill<<<linK, 256, 0, sd>>>(b, r.tl.sortedCell, (uint32_t)KPP, (uint32_t)P);
      k_fill_u32<<<1, 64, 0, sd>>>(s.largeCells, 1, 0u);
      k_splat_order<<<linK, 256, 0, sd>>>(b, (uint32_t)KPP, s.largeCells + 1, s.largeCells);
      k_splat_order_large<<<kSortLargeBlocks, 256, 0, sd>>>(b, s.largeCells + 1, s.largeCells, b.splatRank);
    }
    mark(5);
  } else {
    mark(3);
    mark(4);
    mark(5);
  }
  // The splat entries' shifts follow the sort on the side stream; the current and backup
  // lists, whose flags need the initial candidates and the collect, run on `st`.
  const bool splatsApart = overlap && r.temporal && r.tasks && !r.gather;
  if (overlap) {
    HIP_TRY(hipEventRecord(ctx->joinEv, sd));
    HIP_TRY(hipStreamWaitEvent(st, ctx->joinEv, 0));
  }
  if (splatsApart) {
    k_fill_u32<<<1, 64, 0, sd>>>(r.tl.splatFetch, 1, 0u);
    if (env) k_temporal_tasks<true><<<ctx->taskBlocks[1], 256, 0, sd>>>(a, r.tl, r.kind, kTasksSplats);
    else k_temporal_tasks<false><<<ctx->taskBlocks[0], 256, 0, sd>>>(a, r.tl, r.kind, kTasksSplats);
    HIP_TRY(hipEventRecord(ctx->splatEv, sd));
  }
  if (r.temporal) {
    if (r.gather) {
      if (env) k_gather<true><<<grid, block, 0, st>>>(a);
      else k_gather<false><<<grid, block, 0, st>>>(a);
    } else if (r.tasks) {
      const uint32_t part = splatsApart ? kTasksLists : kTasksAll;
      k_temporal_flags<<<grid, block, 0, st>>>(a, r.tl, r.kind == kTasksBackup ? 1u : 0u);
      if (int32_t e = queue_build(ctx, s.tq, r.tl.q, p, grid, st)) return e;
      if (env) k_temporal_tasks<true><<<ctx->taskBlocks[1], 256, 0, st>>>(a, r.tl, r.kind, part);
      else k_temporal_tasks<false><<<ctx->taskBlocks[0], 256, 0, st>>>(a, r.tl, r.kind, part);
      if (splatsApart) HIP_TRY(hipStreamWaitEvent(st, ctx->splatEv, 0));
      k_temporal_combine<<<grid, block, 0, st>>>(a, r.tl, r.kind);
    }
  }
  mark(6);
  for (const KernelArgs& sa : r.spatial) {
    if (sa.o.spatialNeighbors) {
      k_spatial_flags<<<grid, block, 0, st>>>(sa, r.sp);
      if (int32_t e = queue_build(ctx, s.sq, r.sp.q, p, grid, st)) return e;
      if (env) k_spatial_tasks<true><<<ctx->spatialBlocks[1], 256, 0, st>>>(sa, r.sp);
      else k_spatial_tasks<false><<<ctx->spatialBlocks[0], 256, 0, st>>>(sa, r.sp);
    }
    k_spatial_combine<<<grid, block, 0, st>>>(sa, r.sp);
  }
  mark(7);
  k_resolve<<<grid16, block, 0, st>>>(p, b, r.finalR);
  if (r.colorOut) HIP_TRY(hipMemcpyAsync(r.colorOut, b.color, 16 * P, hipMemcpyDeviceToDevice, st));
  if (r.rayCount) k_ray_count_origin<<<1, 64, 0, st>>>((const unsigned long long*)ctx->rays.ptr, r.rayCount);
  if (r.pathLength) k_fill_u32<<<(uint32_t)((P + 255) / 256), 256, 0, st>>>(r.pathLength, P, 0u);
  mark(8);
  HIP_TRY(hipGetLastError());
  mark(9);
  return RS_OK;
}

// One frame (execute, RS/ReservoirSplatting.cpp:399-509) over prepared buffers.
static int32_t render_frame(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, const FramePlan& f,
                            const FrameScratch& s, hipStream_t st) {
  if (int32_t e = frame_task_blocks(ctx)) return e;
  // Stages overlap unless stage timing is on or the frame is being recorded into a graph (see
  // launch_frame).
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (st) HIP_TRY(hipStreamIsCapturing(st, &capture));
  const bool overlap = ctx->overlap && !ctx->timing && ctx->side && capture == hipStreamCaptureStatusNone;
  FrameRecord r;
  if (int32_t e = record_frame(ctx, cam, io, f, s, overlap, r)) return e;
  return launch_frame(ctx, r, f, s, st);
}

int32_t rs_render_frame(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, void* stream) {
  if (!ctx || !cam || !io) return RS_E_INVALID_ARG;
  hipStream_t st = (hipStream_t)stream;
  FramePlan f;
  if (int32_t e = plan_frame(ctx, cam, io, f)) return e;
  if (!ctx->hasScene) {
    if (io->d_color && ensure_device(ctx) == RS_OK) {
      const size_t n = 4 * f.P;
      k_fill_u32<<<(uint32_t)((n + 255) / 256), 256, 0, st>>>((uint32_t*)io->d_color, n, 0u);
      HIP_TRY(hipGetLastError());
    }
    return fail(ctx, RS_E_NO_SCENE, "no scene set; output cleared");
  }
  if (int32_t e = ensure_device(ctx)) return e;
  FrameScratch s;
  if (int32_t e = prepare_frame(ctx, f, s)) return e;
  return render_frame(ctx, cam, io, f, s, st);
}

// ---- frames without allocation (rs_render_frame_noalloc)
//
// Every temporary buffer of such a frame (task lists, candidate slots, scan storage, the
// Robust gather's ring shifts) is carved from `scratch`, a fixed array that each GPU holds
// for the life of the process.  Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, once per device, on the first call there.
// Calls on one GPU share the array, so calls that overlap on one GPU race.
static constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratchLock;
static void* g_scratchPtr[kMaxDevices] = {};

static int32_t scratch_address(rs_ctx* ctx, void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ctx, RS_E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratchLock);
  if (!g_scratchPtr[dev]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(ctx, RS_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratchPtr[dev] = ptr;
  }
  *out = g_scratchPtr[dev];
  return RS_OK;
}

// Scratch pieces laid out in the fixed array, 256-byte aligned.
static int32_t bind_fixed_scratch(rs_ctx* ctx, std::vector<ScratchPiece>& v, char* base, size_t cap) {
  size_t off = 0;
  for (ScratchPiece& q : v) {
    const size_t at = off;
    off += (q.bytes + 255) & ~(size_t)255;
    if (off > cap)
      return fail(ctx, RS_E_SIZE_MISMATCH,
                  "the frame needs more scratch than rs_render_frame_noalloc holds (" +
                      std::to_string(cap >> 20) + " MiB)");
    *q.out = base + at;
  }
  return RS_OK;
}

int32_t rs_render_frame_noalloc(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, void* stream) {
  if (!ctx || !cam || !io) return RS_E_INVALID_ARG;
  FramePlan f;
  if (int32_t e = plan_frame(ctx, cam, io, f)) return e;
  if (!ctx->hasScene) return fail(ctx, RS_E_NO_SCENE, "no scene set");
  if (!ctx->neighborOffsets.ptr || !ctx->rays.ptr)
    return fail(ctx, RS_E_SIZE_MISMATCH, "render a frame with rs_render_frame first: it allocates the context's buffers");
  void* base = nullptr;
  if (int32_t e = scratch_address(ctx, &base)) return e;
  ctx->noAlloc = true;
  FrameScratch s;
  std::vector<ScratchPiece> pieces;
  int32_t r = prepare_state(ctx, f);
  if (r == RS_OK) r = scratch_pieces(ctx, f, s, pieces);
  if (r == RS_OK) r = bind_fixed_scratch(ctx, pieces, (char*)base, kNoAllocScratchBytes);
  if (r == RS_OK) r = render_frame(ctx, cam, io, f, s, (hipStream_t)stream);
  ctx->noAlloc = false;
  return r;
}

int32_t rs_get_frame_flags(const rs_ctx* ctx, rs_frame_flags* out) {
  if (!ctx || !out) return RS_E_INVALID_ARG;
  *out = ctx->flags;
  return RS_OK;
}

int32_t rs_set_frozen(rs_ctx* ctx, int frozen) {
  if (!ctx) return RS_E_INVALID_ARG;
  ctx->sceneFrozen = frozen != 0;
  return RS_OK;
}

// ------------------------------------------------------------------ batched frames
//
// Many independent contexts (scenes, cameras, resolutions, options) rendered as one unit,
// in one of two forms.
//
// The fused form (the default): the contexts are split into `groups` groups of about equal
// pixel counts, and each group renders with one launch per stage for all its contexts (the
// kb_* kernels), on a stream of its own: the per-pixel kernels' blocks cover every context's
// tiles, contexts ordered by kernel instantiation (environment map, ray-count output) and
// largest frame first; the persistent task kernels drain every context's task lists in
// turn; the scans run over all contexts' lists at once.  Groups overlap on the GPU, so one
// group's memory-bound stages run beside another's compute-bound ones.  The launches are
// captured into a graph once per topology (frame sizes, options, instantiations, groups);
// each frame then costs the host one table write and upload per group and one graph
// launch, and which stages each context runs this frame (none of the temporal ones on its
// first frame) is read from the tables on the device.
//
// The stream form (queues 1..8): each context's frame is recorded on one of `queues` side
// streams forked from the caller's stream, the fork/join is captured into a graph every
// frame, and the instantiated graph is updated in place while its topology is unchanged.
//
// Either way every context's buffers are sized before anything is captured, so no
// allocation is ever recorded, and each context's frame is exactly rs_render_frame's.

// One group of the fused form: its contexts, table and launch parameters.
static constexpr uint32_t kFusedStaging = 8;
struct FusedTopology {
  std::vector<uint32_t> order;   // table position -> context
  uint32_t cls[5] = {0, 0, 0, 0, 0};  // positions of instantiation (env, stats) k: [cls[k], cls[k+1])
  std::vector<uint32_t> tile0;   // per position
  uint32_t tiles = 0;
  std::vector<uint32_t> argsBase;  // per position
  uint32_t numArgs = 0;
  std::vector<std::array<uint32_t, kBatchScans>> chunk0, chunks;  // per position
  uint32_t totalChunks[kBatchScans] = {0, 0, 0, 0};
  uint32_t maxIters = 0;
  bool anyRobust = false, anyCollect = false, anySplat = false, anyMulti = false, anyTasks = false, anyGather = false,
       anyNeighbors = false;
  size_t tableBytes = 0, argsOffset = 0;
  DevBuf table, chunkSums;
  // Pinned host copies of the table, used in turn; a copy is rewritten once its upload,
  // kFusedStaging frames earlier, has run.
  void* staging[kFusedStaging] = {};
  hipEvent_t uploaded[kFusedStaging] = {};
  uint32_t next = 0;
  uint32_t robustBlocks[2] = {0, 0}, taskBlocks[2] = {0, 0}, spatialBlocks[2] = {0, 0};
};

// The fused form's topology: its groups and the graph of all their launches.
struct FusedForm {
  std::string key;
  std::vector<FusedTopology> groups;
  hipGraphExec_t exec = nullptr;
};

struct rs_batch {
  std::vector<rs_ctx*> ctxs;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipEvent_t done = nullptr;   // the last batch frame's end, on the stream it was launched on
  hipStream_t main = nullptr;  // recording stream when the caller passes the null stream
  // Stream form: two executable graphs used on alternate frames, so that a frame's graph
  // is updated while the previous frame's may still run.
  hipGraphExec_t execs[2] = {nullptr, nullptr};
  uint32_t parity = 0;
  uint32_t queues = 0;  // 0: the fused form
  uint32_t groups = 8;  // groups of the fused form (16-scene workload: 1 group 50.2 ms, 4 48.4, 8 47.5)
  uint32_t tileShift = 4;
  bool warm = false;  // a batch frame has been rendered
  FusedForm fused;
  std::string err;
};

static int32_t batch_fail(rs_batch* b, int32_t code, const std::string& msg) {
  b->err = msg;
  return code;
}

#define BATCH_HIP(call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return batch_fail(b, RS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static void release_fused(FusedForm& f) {
  if (f.exec) hipGraphExecDestroy(f.exec);
  for (FusedTopology& t : f.groups) {
    for (uint32_t k = 0; k < kFusedStaging; k++) {
      if (t.uploaded[k]) hipEventDestroy(t.uploaded[k]);
      if (t.staging[k]) hipHostFree(t.staging[k]);
    }
    release(t.table);
    release(t.chunkSums);
  }
  f = FusedForm();
}

int32_t rs_batch_create(rs_ctx* const* ctxs, uint32_t n, rs_batch** out) {
  if (!ctxs || !out || n == 0) return RS_E_INVALID_ARG;
  for (uint32_t i = 0; i < n; i++)
    if (!ctxs[i] || !ctxs[i]->hasScene) return RS_E_NO_SCENE;
  rs_batch* b = new rs_batch();
  b->ctxs.assign(ctxs, ctxs + n);
  b->side.resize(kMaxBatchQueues, nullptr);
  b->join.resize(kMaxBatchQueues, nullptr);
  for (uint32_t q = 0; q < kMaxBatchQueues; q++) {
    if (hipStreamCreateWithFlags(&b->side[q], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->join[q], hipEventDisableTiming) != hipSuccess) {
      rs_batch_destroy(b);
      return RS_E_HIP;
    }
  }
  // A blocking stream: work later queued on the legacy null stream waits for it.
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&b->main, hipStreamDefault) != hipSuccess) {
    rs_batch_destroy(b);
    return RS_E_HIP;
  }
  *out = b;
  return RS_OK;
}

void rs_batch_destroy(rs_batch* b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  for (auto& e : b->execs)
    if (e) hipGraphExecDestroy(e);
  release_fused(b->fused);
  for (auto& s : b->side)
    if (s) hipStreamDestroy(s);
  for (auto& e : b->join)
    if (e) hipEventDestroy(e);
  if (b->fork) hipEventDestroy(b->fork);
  if (b->done) hipEventDestroy(b->done);
  if (b->main) hipStreamDestroy(b->main);
  delete b;
}

const char* rs_batch_last_error(const rs_batch* b) { return b ? b->err.c_str() : "null batch"; }

int32_t rs_batch_configure(rs_batch* b, uint32_t queues, uint32_t tileWidth) {
  if (!b) return RS_E_INVALID_ARG;
  if (queues > kMaxBatchQueues) return batch_fail(b, RS_E_INVALID_ARG, "queues must be 0 (fused) or 1..8");
  uint32_t sh = 0;
  while ((1u << sh) < tileWidth) sh++;
  if ((1u << sh) != tileWidth || sh < 3 || sh > 6) return batch_fail(b, RS_E_INVALID_ARG, "tile width must be 8, 16, 32 or 64");
  if (queues != b->queues) {
    for (auto& e : b->execs) {
      if (e) hipGraphExecDestroy(e);
      e = nullptr;
    }
  }
  b->queues = queues;
  b->tileShift = sh;
  return RS_OK;
}

int32_t rs_batch_configure_fused(rs_batch* b, uint32_t groups, uint32_t tileWidth) {
  if (!b) return RS_E_INVALID_ARG;
  if (groups < 1 || groups > kMaxBatchQueues) return batch_fail(b, RS_E_INVALID_ARG, "groups must be 1..8");
  if (int32_t e = rs_batch_configure(b, 0, tileWidth)) return e;
  b->groups = groups;
  return RS_OK;
}

int32_t rs_batch_get_config(const rs_batch* b, uint32_t* queues, uint32_t* groups, uint32_t* tileWidth) {
  if (!b) return RS_E_INVALID_ARG;
  if (queues) *queues = b->queues;
  if (groups) *groups = b->groups;
  if (tileWidth) *tileWidth = 1u << b->tileShift;
  return RS_OK;
}

// Contexts to queues: largest frame first onto the least loaded queue (by pixels).
static std::vector<uint32_t> batch_assign(const rs_batch* b, const rs_frame_io* ios) {
  const uint32_t n = (uint32_t)b->ctxs.size();
  std::vector<uint32_t> order(n), q(n);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return (size_t)ios[x].width * ios[x].height > (size_t)ios[y].width * ios[y].height;
  });
  std::vector<size_t> load(b->queues, 0);
  for (uint32_t i : order) {
    uint32_t best = 0;
    for (uint32_t k = 1; k < b->queues; k++)
      if (load[k] < load[best]) best = k;
    q[i] = best;
    load[best] += (size_t)ios[i].width * ios[i].height;
  }
  return q;
}

// Stream form: every context's frame recorded on a side stream, one graph per frame.
static int32_t batch_render_streams(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios,
                                    std::vector<FramePlan>& plans, std::vector<FrameScratch>& scratch, hipStream_t st) {
  const uint32_t n = (uint32_t)b->ctxs.size();
  const std::vector<uint32_t> q = batch_assign(b, ios);
  for (uint32_t i = 0; i < n; i++)
    if (int32_t e = frame_task_blocks(b->ctxs[i])) return batch_fail(b, e, b->ctxs[i]->err);
  BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  int32_t code = RS_OK;
  std::string msg;
  if (hipEventRecord(b->fork, st) != hipSuccess) code = RS_E_HIP;
  for (uint32_t k = 0; k < b->queues && code == RS_OK; k++)
    if (hipStreamWaitEvent(b->side[k], b->fork, 0) != hipSuccess) code = RS_E_HIP;
  for (uint32_t i = 0; i < n && code == RS_OK; i++) {
    rs_ctx* c = b->ctxs[i];
    const bool timing = c->timing;
    c->timing = false;  // no event records inside the graph
    c->noAlloc = true;
    code = render_frame(c, &cams[i], &ios[i], plans[i], scratch[i], b->side[q[i]]);
    c->noAlloc = false;
    c->timing = timing;
    if (code != RS_OK) msg = std::string("context ") + std::to_string(i) + ": " + c->err;
  }
  for (uint32_t k = 0; k < b->queues && code == RS_OK; k++) {
    if (hipEventRecord(b->join[k], b->side[k]) != hipSuccess || hipStreamWaitEvent(st, b->join[k], 0) != hipSuccess)
      code = RS_E_HIP;
  }
  hipGraph_t graph = nullptr;
  hipError_t ec = hipStreamEndCapture(st, &graph);
  if (code != RS_OK) {
    if (graph) hipGraphDestroy(graph);
    return batch_fail(b, code, msg.empty() ? "recording the batch failed" : msg);
  }
  if (ec != hipSuccess) return batch_fail(b, RS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  hipGraphExec_t& exec = b->execs[b->parity];
  b->parity ^= 1u;
  if (exec) {
    hipGraphNode_t errNode = nullptr;
    hipGraphExecUpdateResult res;
    bool updated = hipGraphExecUpdate(exec, graph, &errNode, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (!updated) {
      (void)hipGetLastError();
      hipGraphExecDestroy(exec);
      exec = nullptr;
    }
  }
  if (!exec) {
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      hipGraphDestroy(graph);
      exec = nullptr;
      return batch_fail(b, RS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
  }
  hipGraphDestroy(graph);
  BATCH_HIP(hipGraphLaunch(exec, st));
  return RS_OK;
}

// Instantiation class of a context in the fused form: (environment map, ray-count output).
static uint32_t fused_class(const rs_ctx* c, const rs_frame_io& io) {
  return (c->env.present ? 2u : 0u) + (io.d_rayCount ? 1u : 0u);
}

// What fixes the fused form's launches: tile width, and per context its frame size,
// partitions, instantiation and the stages its options can run.
static std::string fused_key(const rs_batch* b, const rs_frame_io* ios, const std::vector<FramePlan>& plans) {
  std::string k = std::to_string(b->tileShift) + "/" + std::to_string(b->groups);
  for (size_t i = 0; i < b->ctxs.size(); i++) {
    const rs_ctx* c = b->ctxs[i];
    const FramePlan& f = plans[i];
    char buf[160];
    std::snprintf(buf, sizeof(buf), "|%u,%u,%u,%u,%d%d%d%d%d,%u,%u,%u", f.W, f.H, f.KP, f.mode, f.robust, f.collects,
                  f.tasks, f.backup, f.spatial, f.neighbors, f.spatial ? c->opt.numSpatialIterations : 0u,
                  fused_class(c, ios[i]));
    k += buf;
  }
  return k;
}

// Occupancy grid of a persistent kernel.
static int32_t fused_blocks(rs_batch* b, const void* kernel, uint32_t& blocks) {
  int perCU = 0, cus = 0, dev = 0;
  BATCH_HIP(hipGetDevice(&dev));
  BATCH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, 256, 0));
  BATCH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  blocks = (uint32_t)std::max(1, perCU * cus);
  return RS_OK;
}

static constexpr uint32_t kBatchLinearBlocks = 512;  // blocks per context of the (blocks, contexts) kernels

// Captures the fused form's launches for the topology `t` on stream st.
static void fused_launches(const FusedTopology& t, hipStream_t st) {
  const BatchTab T = (BatchTab)t.table.ptr;
  const uint32_t n = (uint32_t)t.order.size();
  const dim3 lin(kBatchLinearBlocks, n), few(64, n);
  auto pix = [&](uint32_t c0, uint32_t c1) -> uint32_t {  // tiles of positions c0 .. c1-1
    return c0 < c1 ? (c1 < n ? t.tile0[c1] : t.tiles) - t.tile0[c0] : 0u;
  };
  const uint32_t env0 = t.cls[0], env1 = t.cls[2], envEnd = t.cls[4];
  auto scan = [&](uint32_t kind) {
    if (!t.totalChunks[kind]) return;
    kb_scan_reduce<<<t.totalChunks[kind], 256, 0, st>>>(T, kind);
    kb_scan_chunks<<<n, 256, 0, st>>>(T, kind);
    kb_scan_apply<<<t.totalChunks[kind], 256, 0, st>>>(T, kind);
  };

  kb_begin<<<lin, 256, 0, st>>>(T);
  kb_vbuffer<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
  kb_copy<<<lin, 256, 0, st>>>(T, 0, 1);
  for (uint32_t k = 0; k < 4; k++) {
    const uint32_t c0 = t.cls[k], c1 = t.cls[k + 1], g = pix(c0, c1);
    if (!g) continue;
    switch (k) {
      case 0: kb_initial<false, false><<<g, 256, 0, st>>>(T, c0, c1, t.tile0[c0]); break;
      case 1: kb_initial<false, true><<<g, 256, 0, st>>>(T, c0, c1, t.tile0[c0]); break;
      case 2: kb_initial<true, false><<<g, 256, 0, st>>>(T, c0, c1, t.tile0[c0]); break;
      default: kb_initial<true, true><<<g, 256, 0, st>>>(T, c0, c1, t.tile0[c0]); break;
    }
  }
  kb_copy<<<lin, 256, 0, st>>>(T, 1, 2);
  if (t.anyRobust) {
    kb_robust_flags<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
    scan(kScanRobust);
    if (env0 < env1) kb_robust_tasks<false><<<t.robustBlocks[0], 256, 0, st>>>(T, env0, env1);
    if (env1 < envEnd) kb_robust_tasks<true><<<t.robustBlocks[1], 256, 0, st>>>(T, env1, envEnd);
  }
  if (t.anyCollect) kb_collect<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
  if (t.anySplat) {
    kb_reproject<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
    if (t.anyMulti) {
      if (env0 < env1) kb_multi_reproject<false><<<pix(env0, env1), 256, 0, st>>>(T, env0, env1, t.tile0[env0]);
      if (env1 < envEnd) kb_multi_reproject<true><<<pix(env1, envEnd), 256, 0, st>>>(T, env1, envEnd, t.tile0[env1]);
    }
  }
  if (t.anyTasks) {
    scan(kScanCells);
    kb_splat_fill<<<lin, 256, 0, st>>>(T);
    kb_splat_order<<<lin, 256, 0, st>>>(T);
    kb_splat_order_large<<<few, 256, 0, st>>>(T);
  }
  if (t.anyGather) {
    if (env0 < env1) kb_gather<false><<<pix(env0, env1), 256, 0, st>>>(T, env0, env1, t.tile0[env0]);
    if (env1 < envEnd) kb_gather<true><<<pix(env1, envEnd), 256, 0, st>>>(T, env1, envEnd, t.tile0[env1]);
  }
  if (t.anyTasks) {
    kb_temporal_flags<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
    scan(kScanTemporal);
    if (env0 < env1) kb_temporal_tasks<false><<<t.taskBlocks[0], 256, 0, st>>>(T, env0, env1);
    if (env1 < envEnd) kb_temporal_tasks<true><<<t.taskBlocks[1], 256, 0, st>>>(T, env1, envEnd);
    kb_temporal_combine<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
  }
  for (uint32_t it = 0; it < t.maxIters; it++) {
    if (t.anyNeighbors) {
      kb_spatial_flags<<<t.tiles, 256, 0, st>>>(T, 0, n, 0, it);
      scan(kScanSpatial);
      if (env0 < env1) kb_spatial_tasks<false><<<t.spatialBlocks[0], 256, 0, st>>>(T, env0, env1, it);
      if (env1 < envEnd) kb_spatial_tasks<true><<<t.spatialBlocks[1], 256, 0, st>>>(T, env1, envEnd, it);
    }
    kb_spatial_combine<<<t.tiles, 256, 0, st>>>(T, 0, n, 0, it);
  }
  kb_resolve<<<t.tiles, 256, 0, st>>>(T, 0, n, 0);
  kb_copy<<<lin, 256, 0, st>>>(T, 3, 1);
  kb_end<<<lin, 256, 0, st>>>(T);
}

// One group's topology: context order, tile and chunk spaces, its table.
static int32_t fused_group_build(rs_batch* b, const rs_frame_io* ios, const std::vector<FramePlan>& plans,
                                 std::vector<uint32_t> members, FusedTopology& t) {
  const uint32_t n = (uint32_t)members.size();
  t.order = std::move(members);
  std::stable_sort(t.order.begin(), t.order.end(), [&](uint32_t x, uint32_t y) {
    const uint32_t cx = fused_class(b->ctxs[x], ios[x]), cy = fused_class(b->ctxs[y], ios[y]);
    if (cx != cy) return cx < cy;
    return plans[x].P > plans[y].P;
  });
  for (uint32_t k = 0; k <= 4; k++) {
    uint32_t c = 0;
    while (c < n && fused_class(b->ctxs[t.order[c]], ios[t.order[c]]) < k) c++;
    t.cls[k] = c;
  }
  t.tile0.resize(n);
  t.argsBase.resize(n);
  t.chunk0.resize(n);
  t.chunks.resize(n);
  for (uint32_t k = 0; k < n; k++) {
    const rs_ctx* c = b->ctxs[t.order[k]];
    const FramePlan& f = plans[t.order[k]];
    t.tile0[k] = t.tiles;
    t.tiles += f.grid.x * f.grid.y;
    const uint32_t iters = f.spatial ? c->opt.numSpatialIterations : 0u;
    t.argsBase[k] = t.numArgs;
    t.numArgs += 1 + iters;
    t.maxIters = std::max(t.maxIters, iters);
    const size_t len[kBatchScans] = {f.robust ? (size_t)f.stride : 0, f.tasks ? (f.backup ? 2u : 1u) * (size_t)f.stride : 0,
                                     f.spatial ? 2 * (size_t)f.neighbors * f.stride : 0, f.tasks ? f.KPP : 0};
    for (uint32_t s = 0; s < kBatchScans; s++) {
      t.chunk0[k][s] = t.totalChunks[s];
      t.chunks[k][s] = (uint32_t)((len[s] + kBatchScanChunk - 1) / kBatchScanChunk);
      t.totalChunks[s] += t.chunks[k][s];
    }
    t.anyRobust |= f.robust;
    t.anyCollect |= f.collects;
    t.anySplat |= f.tasks;
    t.anyMulti |= f.mode == RS_TEMPORAL_MULTI_SCATTER;
    t.anyTasks |= f.tasks;
    t.anyGather |= f.collects && f.mode == RS_TEMPORAL_GATHER_ONLY;
    t.anyNeighbors |= f.spatial && f.neighbors > 0;
  }
  t.argsOffset = kBatchHeaderBytes + ((sizeof(BatchCtx) * n + 255) & ~(size_t)255);
  t.tableBytes = t.argsOffset + sizeof(KernelArgs) * t.numArgs;
  uint32_t sums = 0;
  for (uint32_t s = 0; s < kBatchScans; s++) sums += t.totalChunks[s];
  rs_ctx* c0 = b->ctxs[0];
  if (ensure(c0, t.table, t.tableBytes) || ensure(c0, t.chunkSums, 4 * (size_t)std::max(1u, sums)))
    return batch_fail(b, RS_E_OUT_OF_MEMORY, "batch table");
  for (uint32_t k = 0; k < kFusedStaging; k++) {
    BATCH_HIP(hipHostMalloc(&t.staging[k], t.tableBytes, hipHostMallocDefault));
    std::memset(t.staging[k], 0, t.tableBytes);
    BATCH_HIP(hipEventCreateWithFlags(&t.uploaded[k], hipEventDisableTiming));
  }
  if (int32_t e = fused_blocks(b, (const void*)kb_robust_tasks<false>, t.robustBlocks[0])) return e;
  if (int32_t e = fused_blocks(b, (const void*)kb_robust_tasks<true>, t.robustBlocks[1])) return e;
  if (int32_t e = fused_blocks(b, (const void*)kb_temporal_tasks<false>, t.taskBlocks[0])) return e;
  if (int32_t e = fused_blocks(b, (const void*)kb_temporal_tasks<true>, t.taskBlocks[1])) return e;
  if (int32_t e = fused_blocks(b, (const void*)kb_spatial_tasks<false>, t.spatialBlocks[0])) return e;
  return fused_blocks(b, (const void*)kb_spatial_tasks<true>, t.spatialBlocks[1]);
}

// A new topology: the contexts split into groups of about equal pixel counts (largest frame
// first onto the group with the fewest pixels), each group's table, and the graph that runs
// every group's launches on a side stream of its own.
static int32_t fused_build(rs_batch* b, const rs_frame_io* ios, const std::vector<FramePlan>& plans,
                           const std::string& key, hipStream_t st) {
  BATCH_HIP(hipDeviceSynchronize());  // the old topology's tables may still be read
  release_fused(b->fused);
  FusedForm& F = b->fused;
  const uint32_t n = (uint32_t)b->ctxs.size();
  const uint32_t G = std::min(b->groups, n);
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return plans[x].P > plans[y].P; });
  std::vector<std::vector<uint32_t>> members(G);
  std::vector<size_t> load(G, 0);
  for (uint32_t i : order) {
    uint32_t best = 0;
    for (uint32_t g = 1; g < G; g++)
      if (load[g] < load[best]) best = g;
    members[best].push_back(i);
    load[best] += plans[i].P;
  }
  F.groups.resize(G);
  for (uint32_t g = 0; g < G; g++)
    if (int32_t e = fused_group_build(b, ios, plans, std::move(members[g]), F.groups[g])) return e;
  BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
  hipError_t rec = hipSuccess;
  if (G == 1) {
    fused_launches(F.groups[0], st);
  } else {
    rec = hipEventRecord(b->fork, st);
    for (uint32_t g = 0; g < G && rec == hipSuccess; g++) rec = hipStreamWaitEvent(b->side[g], b->fork, 0);
    for (uint32_t g = 0; g < G && rec == hipSuccess; g++) fused_launches(F.groups[g], b->side[g]);
    for (uint32_t g = 0; g < G && rec == hipSuccess; g++) {
      rec = hipEventRecord(b->join[g], b->side[g]);
      if (rec == hipSuccess) rec = hipStreamWaitEvent(st, b->join[g], 0);
    }
  }
  hipGraph_t graph = nullptr;
  hipError_t ec = hipStreamEndCapture(st, &graph);
  if (rec != hipSuccess || ec != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    return batch_fail(b, RS_E_HIP, std::string("recording the fused batch: ") +
                                       hipGetErrorString(rec != hipSuccess ? rec : ec));
  }
  hipError_t ei = hipGraphInstantiate(&F.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    F.exec = nullptr;
    return batch_fail(b, RS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  F.key = key;
  return RS_OK;
}

// One scan segment of a context's frame.
static BatchScan batch_scan(const uint32_t* in, uint32_t* out, uint32_t* queue, uint32_t* fetch, size_t n,
                            uint32_t stride, uint32_t chunk0, uint32_t chunks) {
  return BatchScan{in, out, queue, fetch, (uint32_t)n, stride, chunk0, chunks};
}

// Fused form: each group's table of this frame, uploaded, then the topology's graph.
static int32_t fused_group_table(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios,
                                 std::vector<FramePlan>& plans, std::vector<FrameScratch>& scratch, FusedTopology& t,
                                 hipStream_t st) {
  const uint32_t n = (uint32_t)t.order.size();
  // The staging copy written kFusedStaging frames ago may still be queued for upload.
  BATCH_HIP(hipEventSynchronize(t.uploaded[t.next]));
  unsigned char* tab = (unsigned char*)t.staging[t.next];
  BatchHeader& h = *reinterpret_cast<BatchHeader*>(tab);
  h.n = n;
  h.argsOffset = (uint32_t)t.argsOffset;
  h.chunkSums = (uint32_t*)t.chunkSums.ptr;
  uint32_t base = 0;
  for (uint32_t s = 0; s < kBatchScans; s++) {
    h.chunkBase[s] = base;
    base += t.totalChunks[s];
  }
  BatchCtx* bcs = reinterpret_cast<BatchCtx*>(tab + kBatchHeaderBytes);
  KernelArgs* args = reinterpret_cast<KernelArgs*>(tab + t.argsOffset);
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t i = t.order[k];
    rs_ctx* c = b->ctxs[i];
    const FramePlan& f = plans[i];
    const FrameScratch& s = scratch[i];
    FrameRecord r;
    if (int32_t e = record_frame(c, &cams[i], &ios[i], f, s, false, r))
      return batch_fail(b, e, std::string("context ") + std::to_string(i) + ": " + c->err);
    BatchCtx& bc = bcs[k];
    std::memset(&bc, 0, sizeof(bc));
    const bool neighbors = !r.spatial.empty() && r.a.o.spatialNeighbors > 0;
    bc.stages = (r.robust ? kBRobust : 0u) | (r.collects ? kBCollect : 0u) | (r.splat ? kBSplat : 0u) |
                (r.multi ? kBMulti : 0u) | (r.tasks ? kBTasks : 0u) | (r.gather ? kBGather : 0u) |
                (neighbors ? kBNeighbors : 0u);
    bc.spatialIters = (uint32_t)r.spatial.size();
    bc.args = t.argsBase[k];
    args[bc.args] = r.a;
    for (uint32_t it = 0; it < bc.spatialIters; it++) args[bc.args + 1 + it] = r.spatial[it];
    bc.tile0 = t.tile0[k];
    bc.tiles = f.grid.x * f.grid.y;
    bc.gx = f.grid.x;
    bc.kind = r.kind;
    bc.P = (uint32_t)f.P;
    bc.KPP = (uint32_t)f.KPP;
    bc.tl = r.tl;
    bc.sp = r.sp;
    bc.rq = r.rq;
    bc.largeCells = s.largeCells;
    bc.finalR = r.finalR;
    const auto& c0 = t.chunk0[k];
    const auto& cn = t.chunks[k];
    bc.scan[kScanRobust] = batch_scan(r.rq.flags, r.rq.offsets, r.rq.queue, r.rq.fetch, r.robust ? f.stride : 0,
                                      f.stride, c0[kScanRobust], cn[kScanRobust]);
    bc.scan[kScanTemporal] = batch_scan(r.tl.q.flags, r.tl.q.offsets, r.tl.q.queue, r.tl.q.fetch,
                                        r.tasks ? (size_t)r.tl.q.lists * f.stride : 0, f.stride, c0[kScanTemporal],
                                        cn[kScanTemporal]);
    bc.scan[kScanSpatial] = batch_scan(r.sp.q.flags, r.sp.q.offsets, r.sp.q.queue, r.sp.q.fetch,
                                       neighbors ? (size_t)r.sp.q.lists * f.stride : 0, f.stride, c0[kScanSpatial],
                                       cn[kScanSpatial]);
    bc.scan[kScanCells] = batch_scan(r.a.b.cellCount, r.a.b.cellOffset, nullptr, nullptr, r.tasks ? f.KPP : 0, 0,
                                     c0[kScanCells], cn[kScanCells]);
    const size_t P = f.P;
    if (r.mvecIn) {
      bc.copyDst[0] = (uint2*)r.a.b.mvec;
      bc.copySrc[0] = (const uint2*)r.mvecIn;
      bc.copyN[0] = (uint32_t)P;
    }
    if (r.capture) {
      bc.copyDst[1] = (uint2*)r.captureR;
      bc.copySrc[1] = (const uint2*)r.a.b.currR;
      bc.copyN[1] = (uint32_t)(4 * P);
      bc.copyDst[2] = (uint2*)r.captureC;
      bc.copySrc[2] = (const uint2*)r.a.b.currC;
      bc.copyN[2] = (uint32_t)(14 * P);
    }
    if (r.colorOut) {
      bc.copyDst[3] = (uint2*)r.colorOut;
      bc.copySrc[3] = (const uint2*)r.a.b.color;
      bc.copyN[3] = (uint32_t)(2 * P);
    }
    bc.rayCount = r.rayCount;
    bc.pathLength = r.pathLength;
  }
  BATCH_HIP(hipMemcpyAsync(t.table.ptr, tab, t.tableBytes, hipMemcpyHostToDevice, st));
  BATCH_HIP(hipEventRecord(t.uploaded[t.next], st));
  t.next = (t.next + 1) % kFusedStaging;
  return RS_OK;
}

static int32_t batch_render_fused(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios,
                                  std::vector<FramePlan>& plans, std::vector<FrameScratch>& scratch, hipStream_t st) {
  const std::string key = fused_key(b, ios, plans);
  if (key != b->fused.key || !b->fused.exec)
    if (int32_t e = fused_build(b, ios, plans, key, st)) return e;
  for (FusedTopology& t : b->fused.groups)
    if (int32_t e = fused_group_table(b, cams, ios, plans, scratch, t, st)) return e;
  BATCH_HIP(hipGraphLaunch(b->fused.exec, st));
  return RS_OK;
}

int32_t rs_batch_render(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios, void* stream) {
  if (!b || !cams || !ios) return RS_E_INVALID_ARG;
  // the legacy null stream cannot be captured; record on the batch's own blocking stream
  hipStream_t st = stream ? (hipStream_t)stream : b->main;
  const uint32_t n = (uint32_t)b->ctxs.size();
  for (uint32_t i = 0; i < n; i++) b->ctxs[i]->tileShift = b->tileShift;
  // Every buffer of every frame is sized here, before anything is captured: a new size or
  // an option change allocates now, and the recorded frames find everything in place.
  std::vector<FramePlan> plans(n);
  std::vector<FrameScratch> scratch(n);
  for (uint32_t i = 0; i < n; i++) {
    rs_ctx* c = b->ctxs[i];
    int32_t r = c->hasScene ? plan_frame(c, &cams[i], &ios[i], plans[i]) : fail(c, RS_E_NO_SCENE, "no scene");
    if (r == RS_OK) r = ensure_device(c);
    if (r == RS_OK) r = prepare_frame(c, plans[i], scratch[i]);
    if (r != RS_OK) return batch_fail(b, r, std::string("context ") + std::to_string(i) + ": " + c->err);
  }
  // A batch frame reads what the previous one wrote (histories, the tables): on another
  // stream than the previous frame it first waits for that frame.
  if (b->warm) BATCH_HIP(hipStreamWaitEvent(st, b->done, 0));
  const int32_t r = b->queues == 0 ? batch_render_fused(b, cams, ios, plans, scratch, st)
                                   : batch_render_streams(b, cams, ios, plans, scratch, st);
  if (r == RS_OK) {
    BATCH_HIP(hipEventRecord(b->done, st));
    b->warm = true;
  }
  return r;
}

int32_t rs_batch_autotune(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios, uint32_t frames, void* stream,
                          uint32_t* bestQueues, uint32_t* bestTileWidth, float* bestMs) {
  if (!b || !cams || !ios || frames == 0) return RS_E_INVALID_ARG;
  hipStream_t st = stream ? (hipStream_t)stream : b->main;
  // (queues, groups): the fused form with 1..8 groups, then the stream form with 1..8 queues
  static const uint32_t kForms[][2] = {{0, 1}, {0, 2}, {0, 4}, {0, 6}, {0, 8}, {1, 0}, {2, 0}, {4, 0}, {8, 0}};
  static const uint32_t kTiles[] = {8, 16, 32, 64};
  hipEvent_t e0, e1;
  BATCH_HIP(hipEventCreate(&e0));
  BATCH_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  uint32_t bq = b->queues, bg = b->groups, bt = 1u << b->tileShift;
  if (!b->warm) {
    int32_t r = rs_batch_render(b, cams, ios, stream);
    if (r != RS_OK) return r;
  }
  for (const auto& form : kForms) {
    for (uint32_t tw : kTiles) {
      int32_t r = form[0] ? rs_batch_configure(b, form[0], tw) : rs_batch_configure_fused(b, form[1], tw);
      if (r != RS_OK) return r;
      // one untimed frame records and instantiates this candidate's graph
      if (int32_t r = rs_batch_render(b, cams, ios, stream)) return r;
      BATCH_HIP(hipEventRecord(e0, st));
      for (uint32_t f = 0; f < frames; f++)
        if (int32_t r = rs_batch_render(b, cams, ios, stream)) return r;
      BATCH_HIP(hipEventRecord(e1, st));
      BATCH_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      BATCH_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms /= (float)frames;
      if (ms < best) {
        best = ms;
        bq = form[0];
        bg = form[0] ? b->groups : form[1];
        bt = tw;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->groups = bg;
  if (int32_t r = rs_batch_configure(b, bq, bt)) return r;
  if (bestQueues) *bestQueues = bq;
  if (bestTileWidth) *bestTileWidth = bt;
  if (bestMs) *bestMs = best;
  return RS_OK;
}

int32_t rs_buffer_ptr(rs_ctx* ctx, int which, void** dptr, size_t* bytes) {
  if (!ctx || !dptr || !bytes) return RS_E_INVALID_ARG;
  size_t P = (size_t)ctx->W * ctx->H;
  DevBuf* b = nullptr;
  size_t n = 0;
  switch (which) {
    case RS_BUF_HISTORY_RESERVOIRS: b = &ctx->R[ctx->prev]; n = 32 * P; break;
    case RS_BUF_HISTORY_RECON: b = &ctx->C[ctx->prev]; n = 112 * P; break;
    case RS_BUF_TEMPORAL_RESERVOIRS: b = &ctx->R[ctx->lastTemporal]; n = 32 * P; break;
    case RS_BUF_TEMPORAL_RECON: b = &ctx->C[ctx->lastTemporal]; n = 112 * P; break;
    case RS_BUF_FINAL_RESERVOIRS: b = &ctx->R[ctx->lastFinal]; n = 32 * P; break;
    case RS_BUF_FINAL_RECON: b = &ctx->C[ctx->lastFinal]; n = 112 * P; break;
    case RS_BUF_INITIAL_RESERVOIRS: b = &ctx->R[3]; n = 32 * P; break;
    case RS_BUF_INITIAL_RECON: b = &ctx->C[3]; n = 112 * P; break;
    case RS_BUF_BACKUP_RESERVOIRS: b = &ctx->R[2]; n = 32 * P; break;
    case RS_BUF_BACKUP_RECON: b = &ctx->C[2]; n = 112 * P; break;
    case RS_BUF_MVEC: b = &ctx->mvec; n = 8 * P; break;
    case RS_BUF_VBUFFER: b = &ctx->vbuffer; n = 12 * P; break;
    case RS_BUF_FLOATING_COORDS: b = &ctx->floatingCoords; n = 8 * P; break;
    case RS_BUF_SPLAT_DST: b = &ctx->splatDst; n = 4 * P * ctx->splatPartitions; break;
    case RS_BUF_CELL_COUNT: b = &ctx->cellCount; n = 4 * P * ctx->splatPartitions; break;
    case RS_BUF_CELL_OFFSET: b = &ctx->cellOffset; n = 4 * P * ctx->splatPartitions; break;
    case RS_BUF_SORTED: b = &ctx->sorted; n = 4 * P * ctx->splatPartitions; break;
    case RS_BUF_COLOR: b = &ctx->color; n = 16 * P; break;
    default: return fail(ctx, RS_E_INVALID_ARG, "unknown buffer");
  }
  if (!b->ptr) return fail(ctx, RS_E_INVALID_ARG, "buffer not allocated");
  *dptr = b->ptr;
  *bytes = n;
  return RS_OK;
}

int32_t rs_scatter_counts(rs_ctx* ctx, float* host, size_t* count) {
  if (!ctx || !host || !count) return RS_E_INVALID_ARG;
  size_t P = (size_t)ctx->W * ctx->H;
  if (P == 0) return fail(ctx, RS_E_NO_SCENE, "no frame rendered yet");
  if (*count < P) return fail(ctx, RS_E_SIZE_MISMATCH, "host buffer too small");
  if (!ctx->splatValid) {
    std::memset(host, 0, 4 * P);
  } else {
    if (ensure(ctx, ctx->scatterCounts, 4 * P)) return RS_E_OUT_OF_MEMORY;
    HIP_TRY(hipDeviceSynchronize());
    k_scatter_counts<<<(uint32_t)((P + 255) / 256), 256>>>((const uint32_t*)ctx->cellCount.ptr, (uint32_t)P,
                                                           ctx->splatPartitions, (float*)ctx->scatterCounts.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(host, ctx->scatterCounts.ptr, 4 * P, hipMemcpyDeviceToHost));
  }
  *count = P;
  return RS_OK;
}

static ReplayArgs replay_args(rs_ctx* ctx, const float* params) {
  ReplayArgs a;
  std::memset(&a, 0, sizeof(a));
  a.p = ctx->frozenFrame;
  a.p.W = ctx->frozenW;
  a.p.H = ctx->frozenH;
  fill_scene_view(ctx, a.s);
  a.C = (const PackedRecon*)ctx->frozenC.ptr;
  a.frozen = (FrozenPixel*)ctx->frozen.ptr;
  a.params = params;
  a.numMaterials = ctx->frozenMaterials;
  return a;
}

int32_t rs_freeze(rs_ctx* ctx) {
  if (!ctx) return RS_E_INVALID_ARG;
  if (!ctx->hasScene || !ctx->rendered) return fail(ctx, RS_E_NO_SCENE, "render a frame before freezing it");
  const size_t P = (size_t)ctx->W * ctx->H;
  if (ensure(ctx, ctx->frozen, sizeof(FrozenPixel) * P) || ensure(ctx, ctx->frozenC, 112 * P) ||
      ensure(ctx, ctx->frozenParams0, 16 * (size_t)ctx->numMaterials))
    return RS_E_OUT_OF_MEMORY;
  std::vector<float> p0(4 * (size_t)ctx->numMaterials);
  for (uint32_t m = 0; m < ctx->numMaterials; m++) {
    p0[4 * m] = ctx->hMaterials[m].baseColor[0];
    p0[4 * m + 1] = ctx->hMaterials[m].baseColor[1];
    p0[4 * m + 2] = ctx->hMaterials[m].baseColor[2];
    p0[4 * m + 3] = ctx->hMaterials[m].roughness;
  }
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(ctx->frozenParams0.ptr, p0.data(), 4 * p0.size(), hipMemcpyHostToDevice));
  // the frame's final reconnection records (the history after endFrame's swap, or the
  // current slot while the scene is frozen)
  HIP_TRY(hipMemcpy(ctx->frozenC.ptr, ctx->C[ctx->lastFinal].ptr, 112 * P, hipMemcpyDeviceToDevice));
  ctx->frozenW = ctx->W;
  ctx->frozenH = ctx->H;
  ctx->frozenMaterials = ctx->numMaterials;
  ctx->frozenFrame = ctx->lastParams;
  ReplayArgs a = replay_args(ctx, (const float*)ctx->frozenParams0.ptr);
  k_freeze<<<(uint32_t)((P + 255) / 256), 256>>>(a, (const float*)ctx->color.ptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return RS_OK;
}

static int32_t replay_check(rs_ctx* ctx, size_t numParams, size_t imageFloats) {
  if (!ctx->frozenW) return fail(ctx, RS_E_NO_SCENE, "nothing frozen (call rs_freeze after a frame)");
  if (numParams != 4 * (size_t)ctx->frozenMaterials)
    return fail(ctx, RS_E_SIZE_MISMATCH, "parameters: 4 floats per material of the frozen scene expected");
  if (imageFloats != 4 * (size_t)ctx->frozenW * ctx->frozenH)
    return fail(ctx, RS_E_SIZE_MISMATCH, "image: 4 floats per pixel of the frozen frame expected");
  return RS_OK;
}

int32_t rs_replay_forward(rs_ctx* ctx, const float* d_params, size_t numParams, float* d_image, size_t imageFloats,
                          void* stream) {
  if (!ctx || !d_params || !d_image) return RS_E_INVALID_ARG;
  if (int32_t e = replay_check(ctx, numParams, imageFloats)) return e;
  const size_t P = (size_t)ctx->frozenW * ctx->frozenH;
  ReplayArgs a = replay_args(ctx, d_params);
  k_replay_forward<<<(uint32_t)((P + 255) / 256), 256, 0, (hipStream_t)stream>>>(a, d_image);
  HIP_TRY(hipGetLastError());
  return RS_OK;
}

int32_t rs_replay_backward(rs_ctx* ctx, const float* d_params, size_t numParams, const float* d_grad_image,
                           size_t imageFloats, double* d_grad_params, void* stream) {
  if (!ctx || !d_params || !d_grad_image || !d_grad_params) return RS_E_INVALID_ARG;
  if (int32_t e = replay_check(ctx, numParams, imageFloats)) return e;
  const size_t P = (size_t)ctx->frozenW * ctx->frozenH;
  hipStream_t st = (hipStream_t)stream;
  uint32_t ng = 4 * ctx->frozenMaterials;
  k_zero_f64<<<(ng + 255) / 256, 256, 0, st>>>(d_grad_params, ng);
  ReplayArgs a = replay_args(ctx, d_params);
  k_replay_backward<<<(uint32_t)((P + 255) / 256), 256, 0, st>>>(a, d_grad_image, d_grad_params);
  HIP_TRY(hipGetLastError());
  return RS_OK;
}

// Checkpoint blob: this header, then the history reservoirs (32 B per pixel) and
// reconnection records (112 B per pixel).
struct StateHeader {
  uint32_t magic, version;
  uint32_t W, H;
  uint32_t frameCount, seed, prevNumKeyframes, resetPending;
  uint32_t numTriangles, numMaterials, recordBytes, _pad;
  CameraManager cm;
};
static constexpr uint32_t kStateMagic = 0x54535352u;  // "RSST"
static constexpr uint32_t kStateVersion = 1;

int32_t rs_save_state(rs_ctx* ctx, void* host, size_t* bytes) {
  if (!ctx || !bytes) return RS_E_INVALID_ARG;
  const size_t P = (size_t)ctx->W * ctx->H;
  if (!ctx->hasScene || P == 0 || !ctx->R[ctx->prev].ptr) return fail(ctx, RS_E_NO_SCENE, "no frame rendered yet");
  const size_t need = sizeof(StateHeader) + 144 * P;
  if (!host) {
    *bytes = need;
    return RS_OK;
  }
  if (*bytes < need) return fail(ctx, RS_E_SIZE_MISMATCH, "host buffer too small");
  StateHeader h{};
  h.magic = kStateMagic;
  h.version = kStateVersion;
  h.W = ctx->W;
  h.H = ctx->H;
  h.frameCount = ctx->frameCount;
  h.seed = ctx->seed;
  h.prevNumKeyframes = ctx->prevNumKeyframes;
  h.resetPending = ctx->resetPending ? 1u : 0u;
  h.numTriangles = ctx->numTris;
  h.numMaterials = ctx->numMaterials;
  h.recordBytes = 144;
  h.cm = ctx->cm;
  char* out = (char*)host;
  std::memcpy(out, &h, sizeof(h));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(out + sizeof(h), ctx->R[ctx->prev].ptr, 32 * P, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(out + sizeof(h) + 32 * P, ctx->C[ctx->prev].ptr, 112 * P, hipMemcpyDeviceToHost));
  *bytes = need;
  return RS_OK;
}

int32_t rs_load_state(rs_ctx* ctx, const void* host, size_t bytes) {
  if (!ctx || !host) return RS_E_INVALID_ARG;
  if (bytes < sizeof(StateHeader)) return fail(ctx, RS_E_SIZE_MISMATCH, "state blob too short");
  StateHeader h;
  std::memcpy(&h, host, sizeof(h));
  if (h.magic != kStateMagic || h.version != kStateVersion || h.recordBytes != 144)
    return fail(ctx, RS_E_INVALID_ARG, "not a state blob of this library version");
  if (h.W == 0 || h.H == 0 || h.W > 4096 || h.H > 4096) return fail(ctx, RS_E_INVALID_ARG, "bad frame size in state blob");
  const size_t P = (size_t)h.W * h.H;
  if (bytes != sizeof(StateHeader) + 144 * P) return fail(ctx, RS_E_SIZE_MISMATCH, "state blob size does not match its frame");
  if (!ctx->hasScene || h.numTriangles != ctx->numTris || h.numMaterials != ctx->numMaterials)
    return fail(ctx, RS_E_NO_SCENE, "the state was saved with another scene");
  if (h.prevNumKeyframes > (uint32_t)kMaxCameraKeyframes || h.cm.numKeyframes > (uint32_t)kMaxCameraKeyframes)
    return fail(ctx, RS_E_INVALID_ARG, "bad keyframe count in state blob");
  if (int32_t e = ensure_device(ctx)) return e;
  if (ensure(ctx, ctx->R[ctx->prev], 32 * P) || ensure(ctx, ctx->C[ctx->prev], 112 * P)) return RS_E_OUT_OF_MEMORY;
  const char* in = (const char*)host;
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(ctx->R[ctx->prev].ptr, in + sizeof(h), 32 * P, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ctx->C[ctx->prev].ptr, in + sizeof(h) + 32 * P, 112 * P, hipMemcpyHostToDevice));
  ctx->W = h.W;
  ctx->H = h.H;
  ctx->frameCount = h.frameCount;
  ctx->seed = h.seed;
  ctx->prevNumKeyframes = h.prevNumKeyframes;
  ctx->resetPending = h.resetPending != 0;
  ctx->cm = h.cm;
  ctx->lastFinal = ctx->prev;
  return RS_OK;
}

int32_t rs_read_buffer(rs_ctx* ctx, int which, void* host, size_t* bytes) {
  if (!ctx || !host || !bytes) return RS_E_INVALID_ARG;
  void* d = nullptr;
  size_t n = 0;
  int32_t r = rs_buffer_ptr(ctx, which, &d, &n);
  if (r != RS_OK) return r;
  if (*bytes < n) return fail(ctx, RS_E_SIZE_MISMATCH, "host buffer too small");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(host, d, n, hipMemcpyDeviceToHost));
  *bytes = n;
  return RS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ temporal accumulation
//
// AccumulatePass host logic (Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:130-290):
// overflow handling, reset on frame-size change, the disabled blit, the frame count update
// and one kernel launch per frame.

struct rs_accum {
  rs_accum_options opt{};
  std::string err;
  uint32_t frameCount = 0;
  uint32_t W = 0, H = 0;
  DevBuf sum, corr, sumD;
};

static int32_t fail(rs_accum* a, int32_t code, const std::string& msg) {
  a->err = msg;
  return code;
}

void rs_accum_default_options(rs_accum_options* out) {
  if (!out) return;
  out->enabled = 1;
  out->precision = RS_ACCUM_SINGLE;
  out->maxFrameCount = 0;
  out->overflowMode = RS_ACCUM_OVERFLOW_STOP;
}

static int32_t accum_check(rs_accum* a, const rs_accum_options* o) {
  if (o->precision > RS_ACCUM_DOUBLE) {
    a->err = "precisionMode must be Single, SingleCompensated or Double";
    return RS_E_BAD_OPTION;
  }
  if (o->overflowMode > RS_ACCUM_OVERFLOW_EMA) {
    a->err = "overflowMode must be Stop, Reset or EMA";
    return RS_E_BAD_OPTION;
  }
  return RS_OK;
}

int32_t rs_accum_create(const rs_accum_options* opts, rs_accum** out) {
  if (!out) return RS_E_INVALID_ARG;
  *out = nullptr;
  rs_accum* a = new rs_accum();
  rs_accum_default_options(&a->opt);
  if (opts) {
    if (int32_t e = accum_check(a, opts)) {
      delete a;
      return e;
    }
    a->opt = *opts;
  }
  *out = a;
  return RS_OK;
}

void rs_accum_destroy(rs_accum* a) {
  if (!a) return;
  if (a->sum.ptr || a->corr.ptr || a->sumD.ptr) (void)hipDeviceSynchronize();
  release(a->sum);
  release(a->corr);
  release(a->sumD);
  delete a;
}

const char* rs_accum_last_error(const rs_accum* a) { return a ? a->err.c_str() : "null accumulator"; }

int32_t rs_accum_set_options(rs_accum* a, const rs_accum_options* opts) {
  if (!a || !opts) return RS_E_INVALID_ARG;
  if (int32_t e = accum_check(a, opts)) return e;
  a->opt = *opts;
  a->frameCount = 0;  // mode, limit, overflow and enable changes reset the accumulation
  return RS_OK;
}

int32_t rs_accum_reset(rs_accum* a) {
  if (!a) return RS_E_INVALID_ARG;
  a->frameCount = 0;
  return RS_OK;
}

uint32_t rs_accum_frame_count(const rs_accum* a) { return a ? a->frameCount : 0u; }

int32_t rs_accum_execute(rs_accum* a, const float* d_src, float* d_dst, uint32_t width, uint32_t height,
                         int shouldAccumulate, int viewAccumulation, void* stream) {
  if (!a || !d_src || !d_dst || width == 0 || height == 0) return RS_E_INVALID_ARG;
  hipStream_t st = (hipStream_t)stream;
  rs_accum* ctx = a;  // for HIP_TRY
  const rs_accum_options& o = a->opt;
  // Overflow at the frame limit.
  if (o.maxFrameCount > 0 && a->frameCount == o.maxFrameCount) {
    if (o.overflowMode == RS_ACCUM_OVERFLOW_STOP) return RS_OK;
    if (o.overflowMode == RS_ACCUM_OVERFLOW_RESET) a->frameCount = 0;
  }
  if (width != a->W || height != a->H) {
    a->W = width;
    a->H = height;
    a->frameCount = 0;
  }
  const size_t P = (size_t)width * height;
  if (P > 0xffffffffull) return RS_E_FRAME_TOO_LARGE;
  if (!o.enabled) {
    HIP_TRY(hipMemcpyAsync(d_dst, d_src, 16 * P, hipMemcpyDeviceToDevice, st));
    return RS_OK;
  }
  // Buffers of the precision mode; a new allocation starts the accumulation over.
  auto prepare = [&](DevBuf& b, size_t bytes, bool used) -> int32_t {
    if (!used) {
      release(b);
      return RS_OK;
    }
    if (b.ptr && b.bytes == bytes) return RS_OK;
    release(b);
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
      a->err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return RS_E_OUT_OF_MEMORY;
    }
    b.bytes = bytes;
    a->frameCount = 0;
    return RS_OK;
  };
  const uint32_t prec = o.precision;
  if (int32_t e = prepare(a->sum, 16 * P, prec != RS_ACCUM_DOUBLE)) return e;
  if (int32_t e = prepare(a->corr, 16 * P, prec == RS_ACCUM_SINGLE_COMPENSATED)) return e;
  if (int32_t e = prepare(a->sumD, 32 * P, prec == RS_ACCUM_DOUBLE)) return e;
  AccumArgs k{};
  k.src = (const float4*)d_src;
  k.dst = (float4*)d_dst;
  k.sum = (float4*)a->sum.ptr;
  k.corr = (float4*)a->corr.ptr;
  k.sumD = (double2*)a->sumD.ptr;
  k.pixels = (uint32_t)P;
  k.count = a->frameCount;
  k.accumulate = shouldAccumulate ? 1u : 0u;
  k.view = viewAccumulation ? 1u : 0u;
  k.movingAverage = o.maxFrameCount > 0 ? 1u : 0u;
  if (o.maxFrameCount == 0 || prec == RS_ACCUM_SINGLE_COMPENSATED || a->frameCount < o.maxFrameCount)
    a->frameCount += shouldAccumulate ? 1u : 0u;
  const uint32_t blocks = (uint32_t)((P + 255) / 256);
  if (prec == RS_ACCUM_SINGLE) k_accum_single<<<blocks, 256, 0, st>>>(k);
  else if (prec == RS_ACCUM_SINGLE_COMPENSATED) k_accum_compensated<<<blocks, 256, 0, st>>>(k);
  else k_accum_double<<<blocks, 256, 0, st>>>(k);
  HIP_TRY(hipGetLastError());
  return RS_OK;
}
