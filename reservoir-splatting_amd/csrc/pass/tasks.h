// Candidate-compacted temporal resampling (ScatterOnly, ScatterBackup, MultiScatter).
//
// The reference evaluates each pixel's temporal candidates in one shader invocation
// (RS/ScatterTemporalResampling.rt.slang:65-174, ScatterBackupTemporalResampling.rt.slang:
// 69-283, MultiScatterTemporalResampling.rt.slang:65-186): the current sample, the backup
// sample (ScatterBackup) and every splat that landed on the pixel, each shifted once or
// twice before its resampling weight is known.  A shift is only evaluated for a candidate
// whose integrand is non-zero, and at 1 spp many are zero (about 11% of current and half of
// the backup samples are live in the 1080p field frame), while splat counts differ from
// pixel to pixel.  One lane per pixel therefore runs the shift code with part of a wave's lanes.
//
// Here the same program is split in three so that shifts run on full waves:
//   1. enqueue (per pixel): the pixel's live current and backup candidates are appended to
//      two task lists; every splat-list entry is a live task by itself.
//   2. tasks (persistent kernel; each wave takes 64 tasks at a time): a task evaluates one
//      candidate's shifts and MIS weight, statement for statement as the per-pixel program
//      does, and stores them in the candidate's slot.
//   3. combine (per pixel): the reservoir merges the candidates in the reference's order
//      from the stored values, with the same arithmetic, and the selected candidate's
//      reconnection record is reloaded and updated with its stored shift.
// A candidate's shifts take the pixel's sample generator as it stands before that
// candidate's merge, and every merge draws exactly one number, so the candidate merged
// r-th (0-based) sees the generator advanced r times from the pixel's seed.  The result
// is bit-identical to the per-pixel program whatever order the tasks run in.
#pragma once
#include "stages.h"

namespace rs {

// Slot of one candidate's stored shift (80 B): a 32-byte head read for every candidate
// by the combine, a 48-byte tail read for the selected one.
//   head: [0] mis  [1] jacobian  [2..4] pHat  [5..6] fractionalPixel  [7] primaryHit.prim
//   tail: [8..9] primaryHit.uv  [10..11] lensSample  [12..14] firstRayDir
//         [15] subPixelJacobian  [16] lensVertexJacobian  [17] secondaryPathJacobian  [18..19] 0
struct PackedCand {
  uint32_t w[20];
};

// A compacted task list: the live entries of `lists` lists, each list one entry per
// tile-order position of the frame (the lane order of the per-pixel kernels), in list
// order and tile order within a list.  Built without atomics: flags, an exclusive scan,
// then each live entry writes its pixel index at its scanned position.
struct TaskQueue {
  uint32_t* flags;    // lists x stride: 1 when the entry is live
  uint32_t* offsets;  // exclusive scan of flags: the entry's queue position
  uint32_t* queue;    // pixel index of every live entry
  uint32_t* fetch;    // next task of the persistent kernel
  uint32_t stride;    // tile-order positions per list (blocks x 256)
  uint32_t lists;
};

// Temporal task lists: list 0 the live current samples, list 1 (ScatterBackup) the live
// backup samples; every splat-list entry is a task after them.
struct TaskLists {
  PackedCand* cand;     // slots: current [0, P), backup [P, 2P), splat entry e at splatBase + e
  TaskQueue q;
  uint32_t* sortedCell; // cell of each splat-list entry
  uint32_t splatBase;   // first splat slot (P without backup slots, 2P with)
  uint32_t* splatFetch; // next splat-list task when the splat entries run in a launch of their own
};

#if defined(__HIPCC__)

__device__ __forceinline__ void st_cand_mis(PackedCand* c, size_t i, float mis) { c[i].w[0] = asuint(mis); }

__device__ __forceinline__ void st_cand(PackedCand* c, size_t i, float mis, float J, vec3 pHat,
                                        const ShiftedPathData& s) {
  uint4* q = reinterpret_cast<uint4*>(c + i);
  q[0] = make_uint4(asuint(mis), asuint(J), asuint(pHat.x), asuint(pHat.y));
  q[1] = make_uint4(asuint(pHat.z), asuint(s.fractionalPixel.x), asuint(s.fractionalPixel.y), s.primaryHit.prim);
  q[2] = make_uint4(asuint(s.primaryHit.u), asuint(s.primaryHit.v), asuint(s.lensSample.x), asuint(s.lensSample.y));
  q[3] = make_uint4(asuint(s.firstRayDir.x), asuint(s.firstRayDir.y), asuint(s.firstRayDir.z),
                    asuint(s.subPixelJacobian));
  q[4] = make_uint4(asuint(s.lensVertexJacobian), asuint(s.secondaryPathJacobian), 0u, 0u);
}

struct CandHead {
  float mis, J;
  vec3 pHat;
  vec2 fp;
};

__device__ __forceinline__ CandHead ld_cand_head(const PackedCand* c, size_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(c + i);
  uint4 a = q[0], b = q[1];
  return CandHead{asfloat(a.x), asfloat(a.y), vec3{asfloat(a.z), asfloat(a.w), asfloat(b.x)},
                  vec2{asfloat(b.y), asfloat(b.z)}};
}

// The stored shift, with the fields recon_update reads.
__device__ __forceinline__ ShiftedPathData ld_cand_shift(const PackedCand* c, size_t i) {
  const uint4* q = reinterpret_cast<const uint4*>(c + i);
  uint4 b = q[1], t0 = q[2], t1 = q[3], t2 = q[4];
  ShiftedPathData s = shifted_default();
  s.fractionalPixel = vec2{asfloat(b.y), asfloat(b.z)};
  s.primaryHit = HitInfo{b.w, asfloat(t0.x), asfloat(t0.y)};
  s.lensSample = vec2{asfloat(t0.z), asfloat(t0.w)};
  s.firstRayDir = vec3{asfloat(t1.x), asfloat(t1.y), asfloat(t1.z)};
  s.subPixelJacobian = asfloat(t1.w);
  s.lensVertexJacobian = asfloat(t2.x);
  s.secondaryPathJacobian = asfloat(t2.y);
  return s;
}

// A block's tile of the frame: column bx and row by of a grid gx tiles wide.  A plain
// launch takes it from the block index; a batched launch, whose blocks cover the tiles of
// many frames, from its tile table.
struct TileId {
  uint32_t bx, by, gx;
};
__device__ __forceinline__ TileId grid_tile() { return TileId{blockIdx.x, blockIdx.y, gridDim.x}; }

// Tile-order position of a lane (blocks of 256 lanes in grid order).
__device__ __forceinline__ uint32_t tile_position(const TileId& t, uint32_t lane) {
  return (t.by * t.gx + t.bx) * 256u + lane;
}
__device__ __forceinline__ uint32_t tile_position(const TileId& t) { return tile_position(t, threadIdx.x); }
__device__ __forceinline__ void queue_flag(const TaskQueue& q, uint32_t list, uint32_t pos, bool live) {
  q.flags[(size_t)list * q.stride + pos] = live ? 1u : 0u;
}
// First queue position of list l (l <= lists; l == lists gives the total).
__device__ __forceinline__ uint32_t queue_start(const TaskQueue& q, uint32_t l) {
  if (l < q.lists) return q.offsets[(size_t)l * q.stride];
  const size_t last = (size_t)q.lists * q.stride - 1;
  return q.offsets[last] + q.flags[last];
}

// Number of splat-list entries of the frame's K partitions.
__device__ __forceinline__ uint32_t splat_total(const PassBuffers& b, uint32_t KP) {
  return b.cellOffset[KP - 1] + b.cellCount[KP - 1];
}

// Merge position of splat entry e among its pixel's candidates (the entries of the
// pixel's cells in partition order), and the pixel and partition.
__device__ __forceinline__ uint32_t splat_rank(const PassBuffers& b, const TaskLists& t, uint32_t P, uint32_t e,
                                               uint32_t& idx, uint32_t& j) {
  const uint32_t cell = t.sortedCell[e];
  j = cell / P;
  idx = cell - j * P;
  uint32_t r = e - b.cellOffset[cell];
  for (uint32_t k = 0; k < j; k++) r += b.cellCount[k * P + idx];
  return r;
}

// ------------------------------------------------------------------ ScatterOnly / MultiScatter

// The current sample's MIS weight (scatter_temporal's and multi_scatter's first block).
__device__ __forceinline__ void scatter_current_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                            const TaskLists& t, bool multi, uint32_t idx) {
  const FrameParams& p = *c.p;
  const bool uc = o.useConfidence != 0;
  const float dt = p.cm.artificialFrameTime;
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_make(x, y, pass_seed(p, 5));
  PathReservoir cr = ldR(b.currR, idx);
  ReconnectionData cc = ldC(b.currC, idx);
  float m1 = luminance(cr.integrand) * conf_or_one(uc, cr.confidence);
  float time = cc.time + dt;
  if (multi) {
    // The current sample moves to the previous frame at its time folded into the first
    // partition, stretched to the whole shutter.
    const float shutter = p.camera.shutterSpeed;
    const float timePartitions = shutter / (float)o.numPartitions;
    float fractionalTime = (cc.time - ffloor(cc.time / timePartitions) * timePartitions) / timePartitions;
    time = fractionalTime * shutter + dt;
  }
  ShiftedPathData sh = scatter_reprojection_shift(c, sg, cc, time, cc.firstHit, cc.lensSample, false);
  float J = scatter_jacobian(sh, cc);
  float m2 = luminance(sh.radiance) * J;
  m2 = fisnan(m2) ? 0.f : m2;
  int sx = (int)ffloor(sh.fractionalPixel.x), sy = (int)ffloor(sh.fractionalPixel.y);
  if (pixel_in(p, sx, sy)) {
    PathReservoir pr = ldR(b.prevR, (uint32_t)sy * p.W + (uint32_t)sx);
    m2 *= conf_or_one(uc, pr.confidence);
  } else {
    m2 = 0.f;
  }
  st_cand_mis(t.cand, idx, (m1 + m2 == 0.f) ? 0.f : m1 / (m1 + m2));
}

// One splat entry's shift to its destination pixel (the splat loop body).
__device__ __forceinline__ void scatter_splat_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                          const TaskLists& t, bool multi, uint32_t e) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidence != 0;
  uint32_t idx, j;
  const uint32_t r = splat_rank(b, t, P, e, idx, j);
  const uint32_t si = b.sorted[e];
  PathReservoir pr = ldR(b.prevR, si);
  if (!any_pos(pr.integrand)) return;  // the combine merges it with weight 0
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_skip(rng_make(x, y, pass_seed(p, 5)), 1u + r);
  ReconnectionData pc = ldC(b.prevC, si);
  float time = pc.time;
  if (multi) {
    const float shutter = p.camera.shutterSpeed;
    time = (pc.time / shutter + (float)j) * (shutter / (float)o.numPartitions);
  }
  // the current sample's confidence is the initial candidates' constant, not read from them,
  // so that the splat entries can be shifted while the initial candidates are traced
  const float crConfidence = kInitialConfidence;
  ShiftedPathData sh = scatter_reprojection_shift(c, sg, pc, time, pc.firstHit, pc.lensSample, false);
  float J = scatter_jacobian(sh, pc);
  float m1 = luminance(sh.radiance) * J * conf_or_one(uc, crConfidence);
  vec3 prevPHat = fisnan(m1) ? v3s(0.f) : sh.radiance;
  J = fisnan(m1) ? 1.f : J;
  float m2 = luminance(pr.integrand) * conf_or_one(uc, pr.confidence);
  float prevMIS = ((m1 + m2) == 0.f) ? 0.f : m2 / (m1 + m2);
  st_cand(t.cand, (size_t)t.splatBase + e, prevMIS, J, prevPHat, sh);
}

// The merge of scatter_temporal / multi_scatter from the stored candidates.
__device__ inline void scatter_combine(const FrameParams& p, const PassBuffers& b, const TemporalOptions& o,
                                       const TaskLists& t, bool multi, uint32_t x, uint32_t y) {
  const uint32_t P = p.W * p.H, K = o.numPartitions;
  const bool uc = o.useConfidence != 0;
  const float jacobianScale = 1.f / (float)K;
  Rng sg = rng_make(x, y, pass_seed(p, 5));
  PathReservoir dst = reservoir_default();
  const uint32_t idx = y * p.W + x;
  PathReservoir cr = ldR(b.currR, idx);
  const float currMIS = any_pos(cr.integrand) ? asfloat(t.cand[idx].w[0]) : 1.f;
  const uint32_t kNone = 0xffffffffu, kCurrent = 0xfffffffeu;
  uint32_t sel = kNone, selJ = 0;
  if (reservoir_add_from(dst, sg, currMIS, cr.integrand, 1.f, cr)) sel = kCurrent;
  float newConfidence = dst.confidence;
  for (uint32_t j = 0; j < K; j++) {
    uint32_t n = b.cellCount[j * P + idx], off = b.cellOffset[j * P + idx];
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t e = off + i;
      PathReservoir pr = ldR(b.prevR, b.sorted[e]);
      float prevMIS = 0.f, J = 1.f;
      vec3 prevPHat = v3s(0.f);
      if (any_pos(pr.integrand)) {
        CandHead h = ld_cand_head(t.cand, (size_t)t.splatBase + e);
        prevMIS = h.mis;
        J = h.J;
        prevPHat = h.pHat;
        pr.subPixel = h.fp - vfloor(h.fp);
      }
      if (reservoir_add_from(dst, sg, prevMIS, prevPHat, multi ? J * jacobianScale : J, pr)) {
        sel = e;
        selJ = j;
      }
    }
  }
  vec2 prevPixel = prev_pixel_from_mvec(p, b, x, y);
  newConfidence = bilinear_confidence(p, b, prevPixel, uc, newConfidence);
  dst.confidence = fmin_(kConfidenceCap, newConfidence);
  ReconnectionData dstC = recon_default();
  if (sel == kCurrent) {
    dstC = ldC(b.currC, idx);
  } else if (sel != kNone) {
    const uint32_t si = b.sorted[sel];
    dstC = ldC(b.prevC, si);
    if (any_pos(ldR(b.prevR, si).integrand)) {
      float oldTime = dstC.time;
      recon_update(dstC, ld_cand_shift(t.cand, (size_t)t.splatBase + sel));
      if (multi) {
        const float shutter = p.camera.shutterSpeed;
        dstC.time = (oldTime / shutter + (float)selJ) * (shutter / (float)K);
      }
    }
  }
  stR(b.currR, idx, dst);
  stC(b.currC, idx, dstC);
}

// ------------------------------------------------------------------ ScatterBackup

__device__ __forceinline__ vec2 floating_coord(const PassBuffers& b, uint32_t idx) {
  return vec2{b.floatingCoords[2 * idx], b.floatingCoords[2 * idx + 1]};
}

// The current sample's MIS weight (scatter_backup's "Current sample" block).
__device__ __forceinline__ void backup_current_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                           const TaskLists& t, uint32_t idx) {
  const FrameParams& p = *c.p;
  const bool uc = o.useConfidence != 0;
  const float dt = p.cm.artificialFrameTime;
  const bool pairwise = o.backupMIS == kBackupPairwise;
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_make(x, y, pass_seed(p, 6));
  PathReservoir cr = ldR(b.currR, idx);
  ReconnectionData cc = ldC(b.currC, idx);
  const float brConfidence = ldR(b.backupR, idx).confidence;
  const vec2 floatingCoord = floating_coord(b, idx);
  float m1 = luminance(cr.integrand) * conf_or_one(uc, cr.confidence);
  float m2 = 0.f;
  ShiftedPathData sc = scatter_reprojection_shift(c, sg, cc, cc.time + dt, cc.firstHit, cc.lensSample, false);
  float sJ = scatter_jacobian(sc, cc);
  m2 = luminance(sc.radiance) * sJ;
  m2 = fisnan(m2) ? 0.f : m2;
  int sx = (int)ffloor(sc.fractionalPixel.x), sy = (int)ffloor(sc.fractionalPixel.y);
  if (pixel_in(p, sx, sy)) {
    PathReservoir pr = ldR(b.prevR, (uint32_t)sy * p.W + (uint32_t)sx);
    m2 *= 0.5f * conf_or_one(uc, pr.confidence);
  } else {
    m2 = 0.f;
  }
  float m3 = 0.f;
  vec2 rel = cr.subPixel;
  if (fpixel_in(p, floatingCoord)) {
    ShiftedPathData sh = gather_lens_vertex_copy_shift(c, sg, cc, cc.time + dt, floatingCoord + rel, cc.lensSample);
    float shJ = sh.secondaryPathJacobian / cc.secondaryPathJacobian;
    m3 = luminance(sh.radiance) * shJ * 0.5f * conf_or_one(uc, brConfidence);
  }
  m3 = fisnan(m3) ? 0.f : m3;
  float currMIS;
  if (!pairwise) {
    currMIS = (m1 + m2 + m3 == 0.f) ? 0.f : m1 / (m1 + m2 + m3);
  } else {
    currMIS = 0.5f * (((m1 + m2 == 0.f) ? 0.f : 0.5f * m1 / (0.5f * m1 + m2)) +
                      ((m1 + m3 == 0.f) ? 0.f : 0.5f * m1 / (0.5f * m1 + m3)));
  }
  st_cand_mis(t.cand, idx, currMIS);
}

// The backup (gathered) sample's shift and MIS weight.
__device__ __forceinline__ void backup_backup_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                          const TaskLists& t, uint32_t idx) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidence != 0;
  const float dt = p.cm.artificialFrameTime;
  const bool pairwise = o.backupMIS == kBackupPairwise;
  const float misScale = pairwise ? 0.5f : 1.f;
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_skip(rng_make(x, y, pass_seed(p, 6)), 1u);
  PathReservoir br = ldR(b.backupR, idx);
  ReconnectionData bc = ldC(b.backupC, idx);
  const float crConfidence = ldR(b.currR, idx).confidence;
  ShiftedPathData sb =
      gather_lens_vertex_copy_shift(c, sg, bc, bc.time, vec2{(float)x, (float)y} + br.subPixel, bc.lensSample);
  float shJ = sb.secondaryPathJacobian / bc.secondaryPathJacobian;
  float m1 = luminance(sb.radiance) * shJ * conf_or_one(uc, crConfidence);
  vec3 backupPHat = fisnan(m1) ? v3s(0.f) : sb.radiance;
  shJ = fisnan(m1) ? 1.f : shJ;
  m1 = fisnan(m1) ? 0.f : m1;
  recon_update(bc, sb);
  float m2 = 0.f;
  if (!pairwise) {
    ShiftedPathData ss = scatter_reprojection_shift(c, sg, bc, bc.time + dt, bc.firstHit, bc.lensSample, false);
    float sJ = scatter_jacobian(ss, bc);
    m2 = luminance(ss.radiance) * sJ * shJ;
    m2 = fisnan(m2) ? 0.f : m2;
    int sx = (int)ffloor(ss.fractionalPixel.x), sy = (int)ffloor(ss.fractionalPixel.y);
    if (pixel_in(p, sx, sy)) {
      PathReservoir pr = ldR(b.prevR, (uint32_t)sy * p.W + (uint32_t)sx);
      m2 *= 0.5f * conf_or_one(uc, pr.confidence);
    } else {
      m2 = 0.f;
    }
  }
  float m3 = luminance(br.integrand) * 0.5f * conf_or_one(uc, br.confidence);
  float backupMIS = misScale * ((m1 + m2 + m3 == 0.f) ? 0.f : m3 / (misScale * m1 + m2 + m3));
  st_cand(t.cand, (size_t)P + idx, backupMIS, shJ, backupPHat, sb);
}

// One splat entry's shifts (scatter_backup's splat loop body).
__device__ __forceinline__ void backup_splat_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                         const TaskLists& t, uint32_t e) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidence != 0;
  const float dt = p.cm.artificialFrameTime;
  const bool pairwise = o.backupMIS == kBackupPairwise;
  const float misScale = pairwise ? 0.5f : 1.f;
  uint32_t idx, j;
  const uint32_t r = splat_rank(b, t, P, e, idx, j);
  const uint32_t si = b.sorted[e];
  PathReservoir pr = ldR(b.prevR, si);
  if (!any_pos(pr.integrand)) return;  // the combine merges it with weight 0
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_skip(rng_make(x, y, pass_seed(p, 6)), 2u + r);
  ReconnectionData pc = ldC(b.prevC, si);
  const float crConfidence = kInitialConfidence;  // as in scatter_splat_task
  const float brConfidence = ldR(b.backupR, idx).confidence;
  const vec2 floatingCoord = floating_coord(b, idx);
  ShiftedPathData sp = scatter_reprojection_shift(c, sg, pc, pc.time, pc.firstHit, pc.lensSample, true);
  float sJ = scatter_jacobian(sp, pc);
  float m1 = luminance(sp.radiance) * sJ * conf_or_one(uc, crConfidence);
  m1 = fisnan(m1) ? 0.f : m1;
  vec3 prevPHat = fisnan(m1) ? v3s(0.f) : sp.radiance;
  sJ = fisnan(m1) ? 1.f : sJ;
  recon_update(pc, sp);
  pr.subPixel = pc.subPixel;
  float m2 = luminance(pr.integrand) * 0.5f * conf_or_one(uc, pr.confidence);
  float m3 = 0.f;
  if (!pairwise) {
    vec2 rel = pr.subPixel;
    if (fpixel_in(p, floatingCoord + rel)) {
      ShiftedPathData sh = gather_lens_vertex_copy_shift(c, sg, pc, pc.time + dt, floatingCoord + rel, pc.lensSample);
      float shJ = sh.secondaryPathJacobian / pc.secondaryPathJacobian;
      m3 = luminance(sh.radiance) * shJ * sJ * 0.5f * conf_or_one(uc, brConfidence);
    }
    m3 = fisnan(m3) ? 0.f : m3;
  }
  float prevMIS = misScale * ((m1 + m2 + m3 == 0.f) ? 0.f : m2 / (misScale * m1 + m2 + m3));
  st_cand(t.cand, (size_t)t.splatBase + e, prevMIS, sJ, prevPHat, sp);
}

// scatter_backup's merge from the stored candidates.
__device__ inline void backup_combine(const FrameParams& p, const PassBuffers& b, const TemporalOptions& o,
                                      const TaskLists& t, uint32_t x, uint32_t y) {
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidence != 0;
  Rng sg = rng_make(x, y, pass_seed(p, 6));
  PathReservoir dst = reservoir_default();
  const uint32_t idx = y * p.W + x;
  const uint32_t kNone = 0xffffffffu, kCurrent = 0xfffffffeu, kBackup = 0xfffffffdu;
  uint32_t sel = kNone;
  vec2 prevPixel = prev_pixel_from_mvec(p, b, x, y);
  {
    PathReservoir cr = ldR(b.currR, idx);
    const float currMIS = any_pos(cr.integrand) ? asfloat(t.cand[idx].w[0]) : 1.f;
    if (reservoir_add_from(dst, sg, currMIS, cr.integrand, 1.f, cr)) sel = kCurrent;
  }
  const PathReservoir br = ldR(b.backupR, idx);
  const bool backupLive = any_pos(br.integrand);
  {
    float backupMIS = 0.f, shJ = 1.f;
    vec3 backupPHat = v3s(0.f);
    if (backupLive) {
      CandHead h = ld_cand_head(t.cand, (size_t)P + idx);
      backupMIS = h.mis;
      shJ = h.J;
      backupPHat = h.pHat;
    }
    if (reservoir_add_from(dst, sg, backupMIS, backupPHat, shJ, br)) sel = kBackup;
  }
  float newConfidence = dst.confidence;
  uint32_t n = b.cellCount[idx], off = b.cellOffset[idx];
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t e = off + i;
    PathReservoir pr = ldR(b.prevR, b.sorted[e]);
    float prevMIS = 0.f, sJ = 1.f;
    vec3 prevPHat = v3s(0.f);
    if (any_pos(pr.integrand)) {
      CandHead h = ld_cand_head(t.cand, (size_t)t.splatBase + e);
      prevMIS = h.mis;
      sJ = h.J;
      prevPHat = h.pHat;
      pr.subPixel = h.fp - vfloor(h.fp);
    }
    if (reservoir_add_from(dst, sg, prevMIS, prevPHat, sJ, pr)) sel = e;
  }
  newConfidence = bilinear_confidence(p, b, prevPixel, uc, newConfidence);
  dst.confidence = fmin_(kConfidenceCap, newConfidence);
  ReconnectionData dstC = recon_default();
  if (sel == kCurrent) {
    dstC = ldC(b.currC, idx);
  } else if (sel == kBackup) {
    dstC = ldC(b.backupC, idx);
    if (backupLive) recon_update(dstC, ld_cand_shift(t.cand, (size_t)P + idx));
  } else if (sel != kNone) {
    const uint32_t si = b.sorted[sel];
    dstC = ldC(b.prevC, si);
    if (any_pos(ldR(b.prevR, si).integrand)) recon_update(dstC, ld_cand_shift(t.cand, (size_t)t.splatBase + sel));
  }
  stR(b.currR, idx, dst);
  stC(b.currC, idx, dstC);
}

#endif

// ------------------------------------------------------------------ spatial reuse
//
// spatial() (RS/SpatialResampling.rt.slang:67-185) in the same three steps.  For each
// neighbour i of a pixel two shifts may run: the central sample into the neighbour (its
// MIS term m1, when the central integrand is non-zero) and the neighbour's sample into the
// pixel (when the neighbour's integrand is non-zero).  Each is a task of its own, in list
// dir * neighbours + i.  Both shifts of neighbour i see the generator after the
// start-index draw and one merge per valid neighbour before i.

struct SpatialTasks {
  float* m1;         // central-into-neighbour MIS term of (i, pixel) at i * P + pixel
  PackedCand* cand;  // neighbour-into-pixel shift of (i, pixel) at i * P + pixel
  TaskQueue q;       // 2 * neighbours lists
};

#if defined(__HIPCC__)

__device__ __forceinline__ uint32_t spatial_seed(const FrameParams& p, const TemporalOptions& o) {
  return p.useFixedSeed ? p.fixedSeed + 7 : 8u * (p.frameCount + p.seed) + 1337u * o.spatialIteration + 7u;
}

// Neighbour i of (x, y) for a start index; false when it falls outside the frame.
__device__ __forceinline__ bool spatial_neighbor(const FrameParams& p, const PassBuffers& b, const TemporalOptions& o,
                                                 uint32_t x, uint32_t y, uint32_t startIdx, uint32_t i, uint32_t& ni) {
  uint32_t oi = (startIdx + i) % kSpatialNeighborSamples;
  vec2 off{(float)b.neighborOffsets[2 * oi] / 127.f, (float)b.neighborOffsets[2 * oi + 1] / 127.f};
  int nx = (int)fround((float)x + o.spatialRadius * off.x);
  int ny = (int)fround((float)y + o.spatialRadius * off.y);
  if (!pixel_in(p, nx, ny)) return false;
  ni = (uint32_t)ny * p.W + (uint32_t)nx;
  return true;
}

// Flags of the pixel's live shifts.
__device__ inline void spatial_flags(const FrameParams& p, const PassBuffers& b, const TemporalOptions& o,
                                     const SpatialTasks& t, bool inside, uint32_t x, uint32_t y, uint32_t pos) {
  const uint32_t n = o.spatialNeighbors;
  const uint32_t idx = inside ? y * p.W + x : 0u;
  Rng sg = rng_make(x, y, spatial_seed(p, o));
  const uint32_t startIdx = (uint32_t)(next1(sg) * (float)kSpatialNeighborSamples);
  const bool centralLive = inside && any_pos(ldR(b.prevR, idx).integrand);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t ni = 0;
    const bool valid = inside && spatial_neighbor(p, b, o, x, y, startIdx, i, ni);
    queue_flag(t.q, i, pos, valid && centralLive);
    queue_flag(t.q, n + i, pos, valid && any_pos(ldR(b.prevR, ni).integrand));
  }
}

// One shift of neighbour i of pixel idx: dir 0 the central sample into the neighbour,
// dir 1 the neighbour's sample into the pixel.
__device__ __forceinline__ void spatial_task(const DevCtx& c, const PassBuffers& b, const TemporalOptions& o,
                                    const SpatialTasks& t, uint32_t dir, uint32_t i, uint32_t idx) {
  const FrameParams& p = *c.p;
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidenceSpatial != 0;
  const uint32_t nCount = o.spatialNeighbors;
  const uint32_t x = idx % p.W, y = idx / p.W;
  Rng sg = rng_make(x, y, spatial_seed(p, o));
  const uint32_t startIdx = (uint32_t)(next1(sg) * (float)kSpatialNeighborSamples);
  uint32_t ni = 0, merged = 0;
  for (uint32_t k = 0; k < i; k++) merged += spatial_neighbor(p, b, o, x, y, startIdx, k, ni) ? 1u : 0u;
  sg = rng_skip(sg, merged);
  spatial_neighbor(p, b, o, x, y, startIdx, i, ni);
  vec2 dofProbs = dof_probabilities(c, b, idx);
  const uint32_t numLensCopy = (uint32_t)fround(dofProbs.x * (float)nCount);
  const bool lensCopy = i < numLensCopy;
  const size_t slot = (size_t)i * P + idx;
  if (dir == 0) {
    PathReservoir cr = ldR(b.prevR, idx);
    ReconnectionData cc = ldC(b.prevC, idx);
    const float nrConfidence = ldR(b.prevR, ni).confidence;
    vec2 fp = vec2{(float)(ni % p.W), (float)(ni / p.W)} + cr.subPixel;
    ShiftedPathData sh = lensCopy ? gather_lens_vertex_copy_shift(c, sg, cc, cc.time, fp, cc.lensSample)
                                  : gather_primary_hit_reconnection_shift(c, sg, cc, cc.time, fp, cc.firstHit);
    float J = (lensCopy ? 1.f : (sh.lensVertexJacobian / cc.lensVertexJacobian)) *
              (sh.secondaryPathJacobian / cc.secondaryPathJacobian);
    float m1 = luminance(sh.radiance) * J * conf_or_one(uc, nrConfidence);
    t.m1[slot] = fisnan(m1) ? 0.f : m1;
  } else {
    PathReservoir nr = ldR(b.prevR, ni);
    ReconnectionData nc = ldC(b.prevC, ni);
    const float crConfidence = ldR(b.prevR, idx).confidence;
    vec2 fp = vec2{(float)x, (float)y} + nr.subPixel;
    ShiftedPathData sh = lensCopy ? gather_lens_vertex_copy_shift(c, sg, nc, nc.time, fp, nc.lensSample)
                                  : gather_primary_hit_reconnection_shift(c, sg, nc, nc.time, fp, nc.firstHit);
    float J = (lensCopy ? 1.f : (sh.lensVertexJacobian / nc.lensVertexJacobian)) *
              (sh.secondaryPathJacobian / nc.secondaryPathJacobian);
    float nW = luminance(sh.radiance) * J * conf_or_one(uc, crConfidence);
    vec3 nPHat = fisnan(nW) ? v3s(0.f) : sh.radiance;
    J = fisnan(nW) ? 1.f : J;
    nW = fisnan(nW) ? 0.f : nW / (float)nCount;
    float m2 = luminance(nr.integrand) * conf_or_one(uc, nr.confidence);
    float nMIS = (m2 + nW > 0.f) ? m2 / (nW + m2) : 0.f;
    st_cand(t.cand, slot, nMIS, J, nPHat, sh);
  }
}

// spatial()'s merge from the stored shifts; writes outR/outC.
__device__ inline void spatial_combine(const FrameParams& p, const PassBuffers& b, const TemporalOptions& o,
                                       const SpatialTasks& t, uint32_t x, uint32_t y) {
  const uint32_t P = p.W * p.H;
  const bool uc = o.useConfidenceSpatial != 0;
  const uint32_t nCount = o.spatialNeighbors;
  Rng sg = rng_make(x, y, spatial_seed(p, o));
  const uint32_t idx = y * p.W + x;
  PathReservoir cr = ldR(b.prevR, idx);
  const bool centralLive = any_pos(cr.integrand);
  float centralMIS = 1.f;
  float centralWeight = luminance(cr.integrand) / (float)nCount * conf_or_one(uc, cr.confidence);
  PathReservoir dst = reservoir_default();
  const uint32_t kNone = 0xffffffffu, kCentral = 0xfffffffeu;
  uint32_t sel = kNone, selNi = 0;
  int validNeighbors = 0;
  const uint32_t startIdx = (uint32_t)(next1(sg) * (float)kSpatialNeighborSamples);
  for (uint32_t i = 0; i < nCount; i++) {
    uint32_t ni;
    if (!spatial_neighbor(p, b, o, x, y, startIdx, i, ni)) continue;
    validNeighbors += 1;
    const size_t slot = (size_t)i * P + idx;
    PathReservoir nr = ldR(b.prevR, ni);
    float m1 = centralLive ? t.m1[slot] : 0.f;
    centralMIS += 1.f;
    centralMIS -= (m1 + centralWeight > 0.f) ? m1 / (m1 + centralWeight) : 0.f;
    float nMIS = 0.f, J = 1.f;
    vec3 nPHat = v3s(0.f);
    if (any_pos(nr.integrand)) {
      CandHead h = ld_cand_head(t.cand, slot);
      nMIS = h.mis;
      J = h.J;
      nPHat = h.pHat;
    }
    if (reservoir_add_from(dst, sg, nMIS, nPHat, J, nr)) {
      sel = i;
      selNi = ni;
    }
  }
  if (reservoir_add_from(dst, sg, centralMIS, cr.integrand, 1.f, cr)) sel = kCentral;
  dst.totalWeight /= (float)(validNeighbors + 1);
  ReconnectionData dstC = recon_default();
  if (sel == kCentral) {
    dstC = ldC(b.prevC, idx);
  } else if (sel != kNone) {
    dstC = ldC(b.prevC, selNi);
    if (any_pos(ldR(b.prevR, selNi).integrand)) recon_update(dstC, ld_cand_shift(t.cand, (size_t)sel * P + idx));
  }
  stR(b.outR, idx, dst);
  stC(b.outC, idx, dstC);
}

#endif

}  // namespace rs
