/* C ABI of the MI355X ReservoirSplatting per-frame pass.
 *
 * This header replaces the Falcor render-pass plugin boundary of the reference:
 *   - extern "C" registerPlugin / createPass("ReservoirSplatting", props)
 *       Source/RenderPasses/ReservoirSplatting/ReservoirSplatting.cpp:136-140,
 *       scripts/ReservoirSplatting.py:25-27
 *   - RenderPass::setProperties / getProperties  ReservoirSplatting.cpp:182-344
 *   - RenderPass::setScene                       ReservoirSplatting.cpp:378
 *   - RenderPass::execute                        ReservoirSplatting.cpp:399-509
 *   - reset()                                    ReservoirSplatting.cpp:813-820
 *   - pixelStats (ray counters)                  ReservoirSplatting.cpp:142-156
 * Plain pointers and sizes only.  Every call returns 0 on success or a negative RS_E_*
 * code; rs_last_error() gives the message.  A context owns its history and scratch
 * device buffers; the caller owns the I/O buffers and the stream.  One context per
 * stream, not thread-safe; independent contexts may run concurrently.
 */
#ifndef RS_AMD_H
#define RS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (mirror the reference's failure cases, ReservoirSplatting.cpp:161-164,
 * 192-301, 362-365, 1571-1592). */
#define RS_OK 0
#define RS_E_INVALID_ARG (-1)
#define RS_E_FRAME_TOO_LARGE (-2)
#define RS_E_SIZE_MISMATCH (-3)
#define RS_E_NO_SCENE (-4)
#define RS_E_BAD_OPTION (-5)
#define RS_E_HIP (-6)
#define RS_E_OUT_OF_MEMORY (-7)
#define RS_E_UNSUPPORTED (-8)

/* Enum values identical to RS/ShiftOptions.slang:34-83 and RS/Params.slang:48-53. */
#define RS_TEMPORAL_GATHER_ONLY 0
#define RS_TEMPORAL_SCATTER_ONLY 1
#define RS_TEMPORAL_SCATTER_BACKUP 2
#define RS_TEMPORAL_MULTI_SCATTER 3
#define RS_GATHER_FAST 0
#define RS_GATHER_CLAMPED 1
#define RS_GATHER_ROBUST 2
#define RS_BACKUP_MIS_BALANCE 0
#define RS_BACKUP_MIS_PAIRWISE 1
#define RS_MIS_BALANCE 0
#define RS_MIS_POWER_TWO 1
#define RS_MIS_POWER_EXP 2

/* Options, field for field with StaticParams / ReSTIRParams / CameraParams
 * (RS/ReservoirSplatting.h:124-184) and PathTracerParams (RS/Params.slang:83-86).
 * rs_default_options() returns the header defaults; the reference graph script
 * (scripts/ReservoirSplatting.py:13-21) switches temporalReuse to ScatterOnly. */
typedef struct rs_options {
  uint32_t samplesPerPixel;        /* 1..16 */
  uint32_t maxSurfaceBounces;      /* <= 254 */
  uint32_t maxDiffuseBounces;
  uint32_t maxSpecularBounces;
  uint32_t maxTransmissionBounces;
  uint32_t useRussianRoulette;
  uint32_t misHeuristic;           /* RS_MIS_* */
  float misPowerExponent;
  float specularRoughnessThreshold; /* [0,1] */
  uint32_t adjustShadingNormals;    /* secondary hits */
  uint32_t gbufferAdjustShadingNormals; /* primary hits (VBuffer dictionary flag) */
  uint32_t enableTemporalResampling;
  uint32_t temporalReuse;           /* RS_TEMPORAL_* */
  uint32_t gatherMechanism;         /* RS_GATHER_* */
  uint32_t scatterBackupMIS;        /* RS_BACKUP_MIS_* */
  uint32_t numTimePartitions;
  uint32_t useConfidenceWeightsTemporally;
  uint32_t enableSpatialResampling;
  uint32_t numSpatialIterations;
  uint32_t neighborCount;
  uint32_t useConfidenceWeightsSpatially;
  float spatialGatherRadius;
  float artificialFrameTime;
  uint32_t useFixedSeed;
  uint32_t fixedSeed;
  uint32_t captureStages;           /* nonzero: keep a copy of the initial candidates */
  uint32_t quirkMask;               /* RS_QUIRK_* bits: reference behaviour reproduced (default all) */
} rs_options;

/* Reference quirks, reproduced while their bit is set in rs_options.quirkMask (the
 * default); clearing a bit gives the corrected behaviour.
 *   RS_QUIRK_ROBUST_GREY       RS/CollectTemporalSamples.cs.slang:234: the Robust collect
 *                              stores luminance(radiance) as a grey integrand; cleared, the
 *                              shifted radiance itself.
 *   RS_QUIRK_PATH_TIME_ZERO    RS/PathTracer.slang:196-221: generatePath sets the sample time
 *                              in the reconnection record but leaves the path's time 0, so the
 *                              primary-hit Jacobians use the camera at t = 0; cleared, the
 *                              sample's time.
 *   RS_QUIRK_MULTI_PROJECTION  RS/MultiReprojectTemporalSamples.rt.slang:105-147: splat targets
 *                              come from a tanFovY projection from the camera position that
 *                              ignores the lens, with an unnormalised visibility ray; cleared,
 *                              the single-splat reprojection through the sample's lens point
 *                              (RS/ReprojectTemporalSamples.rt.slang:85-150) at each time.
 *   RS_QUIRK_MIS_DROPS_LIGHT_HITS  RS/PathTracer.slang:898 and :1460: after a light sample, a
 *                              BSDF-sampled hit on an emissive triangle or the environment is
 *                              not counted, although the light sample carries a balance-
 *                              heuristic weight (Falcor's path tracer counts it under MIS), so
 *                              the image loses the BSDF-sampled share; cleared, the hit is
 *                              counted with its MIS weight.
 *   RS_QUIRK_SPP_LAST_SAMPLE   RS/InitialCandidates.cs.slang:54: addCandidateReservoir starts a
 *                              fresh reservoir when the sample index is samplesPerPixel - 1, the
 *                              last sample of the loop (:137-142), so with spp > 1 the earlier
 *                              samples are discarded and the kept one carries weight 1 / spp:
 *                              the image is 1 / spp of a one-sample estimate; cleared, the
 *                              reservoir starts at sample 0 and merges every sample. */
#define RS_QUIRK_ROBUST_GREY 0x1u
#define RS_QUIRK_PATH_TIME_ZERO 0x2u
#define RS_QUIRK_MULTI_PROJECTION 0x4u
#define RS_QUIRK_MIS_DROPS_LIGHT_HITS 0x8u
#define RS_QUIRK_SPP_LAST_SAMPLE 0x10u
#define RS_QUIRKS_REFERENCE 0x1Fu

/* One MetalRough StandardMaterial without textures. */
typedef struct rs_material {
  float baseColor[3];
  float roughness;
  float emissive[3]; /* emissive color times emissive factor */
  float metallic;
  float ior;
  uint32_t activeLobes; /* LobeType bits; 0xff = all */
  uint32_t doubleSided;
  uint32_t _pad;
} rs_material;

#define RS_DIFFUSE_FROSTBITE 0
#define RS_DIFFUSE_LAMBERT 1

/* One geometry instance of the scene: a contiguous vertex range and a contiguous triangle
 * range of the soup (the triangles index only the instance's own vertices).  The scene's
 * instance list takes the place of Falcor's GeometryInstanceData (Scene.h:132-136,
 * Scene.cpp:3387-3953): its index is the `instanceID` of a PackedHitInfo and
 * rs_update_instances moves its vertices rigidly. */
typedef struct rs_instance_desc {
  uint32_t firstVertex, numVertices;
  uint32_t firstTriangle, numTriangles;
} rs_instance_desc;

/* World-space triangle soup. Host pointers.  `instances` (may be NULL: one instance over the
 * whole soup) splits the soup into instances; each instance's vertices as given here are its
 * object-space vertices, placed with the identity transform. */
typedef struct rs_scene_desc {
  const float* positions; /* 3 * numVertices */
  const float* normals;   /* 3 * numVertices */
  uint32_t numVertices;
  const uint32_t* indices; /* 3 * numTriangles */
  const uint32_t* triMaterial;
  uint32_t numTriangles;
  const rs_material* materials;
  uint32_t numMaterials;
  uint32_t diffuseModel; /* RS_DIFFUSE_* */
  const rs_instance_desc* instances;
  uint32_t numInstances;
} rs_scene_desc;

/* Camera of the current frame: Falcor CameraData fields the pass reads. */
typedef struct rs_camera {
  float position[3];
  float target[3];
  float focalLength;   /* mm */
  float frameHeight;   /* mm */
  float aspectRatio;   /* width / height */
  float focalDistance; /* scene units */
  float apertureRadius;
  float shutterSpeed;  /* seconds */
} rs_camera;

/* Per-frame I/O: device pointers owned by the caller.  The reference pass's inputs `vbuffer`
 * and `mvec` and outputs `color`, `rayCount` and `pathLength`
 * (RS/ReservoirSplatting.cpp:64-96; channel sizes must equal the frame size, :1557-1575).
 *   d_vbuffer   PackedHitInfo, 4 x uint32 per pixel (Source/Falcor/Scene/HitInfo.slang:92-125,
 *               455-487): the hit type in bits 29-31 of word 0 (0 = none, 1 = triangle), then
 *               with vbufferPrimitiveIndexBits = 0 the instance index in the rest of word 0 and
 *               the primitive index in word 1 (the 128-bit header form), or with n > 0 the
 *               instance index in bits n-28 and the primitive index in bits 0..n-1 of word 0
 *               (the 32-bit header form); words 2 and 3 are the barycentrics as float bits.
 *               NULL = the primary hits are traced by the pass (the VBufferRT step).
 *   d_mvec      RG32F motion vectors (prev - curr, normalized); NULL = computed.
 *   d_rayCount  R32Uint: rays traced per pixel (PixelStats ray count, RS/PathTracer.slang:284,
 *               :1424).  Only the initial-candidate kernel sets the logging pixel
 *               (RS/PathTracer.slang:304-308), so the rays of every resampling kernel count at
 *               pixel (0,0), as in the reference.  May be NULL.
 *   d_pathLength R32Uint: the PixelStats path length; the pass never logs one, so the reference
 *               clears it every frame and so does this build.  May be NULL.
 *   sceneFlags  what the pass reads from the Scene each frame: RS_SCENE_SHOULD_RESET
 *               (Scene::shouldReset, RS/ReservoirSplatting.cpp:404), RS_SCENE_REPLAYING
 *               (Scene::isReplaying), RS_SCENE_ACCUMULATION_FRAME and RS_SCENE_VIEW_ACCUMULATION
 *               (Scene::shouldAccumulate / viewAccumulation while replaying,
 *               Source/Falcor/Scene/Scene.cpp:2474-2486). */
#define RS_SCENE_SHOULD_RESET 0x1u
#define RS_SCENE_REPLAYING 0x2u
#define RS_SCENE_ACCUMULATION_FRAME 0x4u
#define RS_SCENE_VIEW_ACCUMULATION 0x8u
typedef struct rs_frame_io {
  uint32_t width, height;
  const float* d_mvec;     /* RG32F motion vectors (prev - curr, normalized); NULL = computed */
  float* d_color;          /* RGBA32F output, width * height * 4 floats; may be NULL */
  const uint32_t* d_vbuffer;
  uint32_t vbufferPrimitiveIndexBits;
  uint32_t sceneFlags;
  uint32_t* d_rayCount;
  uint32_t* d_pathLength;
} rs_frame_io;

/* The render-graph dictionary entries the pass writes after each frame
 * (RS/ReservoirSplatting.cpp:503-506). */
typedef struct rs_frame_flags {
  uint32_t suppressAutoReset; /* kSuppressAutoReset = frozen || replaying */
  uint32_t shouldAccumulate;  /* kAccumulatePassShouldRun */
  uint32_t viewAccumulation;  /* kViewAccumulatingPass */
} rs_frame_flags;

typedef struct rs_stats {
  uint64_t raysClosestHit;  /* logTraceRay(ClosestHit) count of the last frame */
  uint64_t raysVisibility;  /* logTraceRay(Visibility) count of the last frame */
  uint32_t frameCount;
  uint32_t seed;
  float msTotal;            /* device time of the last frame (0 unless timing on) */
  float msStage[12];        /* per stage, see RS_STAGE_* */
  uint32_t bvhNodes;        /* 4-wide BVH nodes of the current scene */
  uint32_t bvhLevels;       /* its depth in levels */
  uint32_t bvhBuild;        /* RS_BVH_BUILD_*: how the current tree was built */
} rs_stats;

/* The clustering build (PLOC with the surface-area collapse to 4-wide nodes), or, when that
 * tree is deeper than the traversal stack holds (three entries per level), a balanced tree
 * over the triangles in Morton order, whose depth grows with log2 of the triangle count. */
#define RS_BVH_BUILD_CLUSTERING 0
#define RS_BVH_BUILD_BALANCED 1

#define RS_STAGE_VBUFFER 0
#define RS_STAGE_INITIAL 1
#define RS_STAGE_COLLECT 2
#define RS_STAGE_REPROJECT 3
#define RS_STAGE_SORT 4
#define RS_STAGE_TEMPORAL 5
#define RS_STAGE_SPATIAL 6
#define RS_STAGE_RESOLVE 7

/* Buffers readable with rs_read_buffer (device -> host copy, synchronous). */
#define RS_BUF_HISTORY_RESERVOIRS 0  /* 32 B / pixel: final reservoirs of the last frame */
#define RS_BUF_HISTORY_RECON 1       /* 112 B / pixel */
#define RS_BUF_TEMPORAL_RESERVOIRS 2 /* input of the last spatial iteration (the final ones without spatial) */
#define RS_BUF_TEMPORAL_RECON 3
#define RS_BUF_INITIAL_RESERVOIRS 4  /* needs captureStages */
#define RS_BUF_INITIAL_RECON 5
#define RS_BUF_BACKUP_RESERVOIRS 6   /* collect output */
#define RS_BUF_BACKUP_RECON 7
#define RS_BUF_MVEC 8                /* 8 B / pixel */
#define RS_BUF_VBUFFER 9             /* 12 B / pixel: prim, u, v */
#define RS_BUF_FLOATING_COORDS 10    /* 8 B / pixel */
#define RS_BUF_SPLAT_DST 11          /* 4 B / pixel */
#define RS_BUF_CELL_COUNT 12
#define RS_BUF_CELL_OFFSET 13
#define RS_BUF_SORTED 14
#define RS_BUF_COLOR 15              /* 16 B / pixel, the context's own copy */
#define RS_BUF_FINAL_RESERVOIRS 16   /* the last frame's final reservoirs: the history, or while the */
#define RS_BUF_FINAL_RECON 17        /* scene is frozen the current slot (the history stays put)   */

typedef struct rs_ctx rs_ctx;

/* Header defaults of RS/ReservoirSplatting.h:124-184 (temporal GatherOnly + Fast). */
void rs_default_options(rs_options* out);

/* createPass("ReservoirSplatting", props): validates and clamps like parseProperties
 * (ReservoirSplatting.cpp:192-301); a clamped option returns RS_E_BAD_OPTION after
 * creating the context with the clamped value. */
int32_t rs_create(const rs_options* opts, rs_ctx** out);
void rs_destroy(rs_ctx* ctx);
const char* rs_last_error(const rs_ctx* ctx);

/* setProperties: same validation; history is kept (call rs_reset to drop it). */
int32_t rs_set_options(rs_ctx* ctx, const rs_options* opts);
int32_t rs_get_options(const rs_ctx* ctx, rs_options* out);

/* setScene: uploads the soup, builds the emissive light collection, the alias table
 * and the LBVH on the GPU. */
int32_t rs_set_scene(rs_ctx* ctx, const rs_scene_desc* scene);

/* Replace vertex positions and normals (same topology), for moving geometry such as
 * rigidly moving instances: the light collection is rebuilt and the BVH is refitted
 * (same tree, new boxes; hits stay exact, traversal cost drifts as geometry moves).
 * This is the reference scene's UpdateMode::Refit for dynamic geometry
 * (Source/Falcor/Scene/Scene.cpp:3387-3700 buildBlas, :3927-3953 buildTlas;
 * Scene.h:132-136).  Host pointers. */
int32_t rs_update_vertices(rs_ctx* ctx, const float* positions, const float* normals, uint32_t numVertices);

/* Environment-map light: the scene's EnvMap (Source/Falcor/Scene/Lights/EnvMap.cpp:62-136,
 * EnvMap.slang:60-118) with its EnvMapSampler (Source/Falcor/Rendering/Lights/
 * EnvMapSampler.cpp:44-114).  `texels` is a host RGBA32F lat-long image of width x height
 * (row 0 at the top; u = 0.5 looks along -z); radiance = texel * intensity * tint, rotated
 * by Euler angles XYZ in degrees.  The importance map (512 x 512 equal-area octahedral
 * texels of mean luminance, with mips) is built on the GPU.  Light sampling then picks the
 * environment map or the emissive triangles with equal probability, and misses see the
 * environment (RS/PathTracer.slang:516-534, 604-701, 1433-1569). */
typedef struct rs_env_map_desc {
  const float* texels;
  uint32_t width, height;
  float intensity;
  float tint[3];
  float rotation[3];
} rs_env_map_desc;
int32_t rs_set_env_map(rs_ctx* ctx, const rs_env_map_desc* desc);
int32_t rs_clear_env_map(rs_ctx* ctx);

/* Full BVH build over the current positions (what rs_set_scene does): the reference's
 * UpdateMode::Rebuild (Scene.h:132-136), for callers that refit every frame and rebuild
 * every few frames. */
int32_t rs_rebuild_bvh(rs_ctx* ctx);

/* Rigid instance motion: the scene's instance transforms (Scene::updateGeometryInstances
 * with the BLAS refit and TLAS rebuild, Source/Falcor/Scene/Scene.cpp:3387, :3927).
 * `xforms` holds n row-major 3x4 matrices (host memory), one per instance of the scene
 * (n = its instance count); instance i's world vertices become xforms[i] applied to its
 * object-space vertices, normals transform by the cofactor matrix and are renormalized.
 * The transform, the emissive-triangle update and the BVH refit run on the GPU; only the
 * light powers return to the host, for the alias table. */
int32_t rs_update_instances(rs_ctx* ctx, const float* xforms, uint32_t n);

/* execute(): one frame of the pass on `stream` (a hipStream_t, NULL = default).  The frame
 * forks the stages that do not need its initial candidates onto a stream the context owns and
 * joins them back into `stream`, so work the caller enqueues after the call sees the whole
 * frame.  With stage timing on, or while `stream` is capturing a graph, every stage runs on
 * `stream` in order. */
int32_t rs_render_frame(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, void* stream);

/* No counterpart in the reference.  rs_render_frame for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same frame, bit for bit, with no hipMalloc, hipFree or allocating library
 * call.  The context's own buffers must already fit the frame (an earlier rs_render_frame at
 * this size or larger, with the same options); otherwise RS_E_SIZE_MISMATCH.  The frame's
 * temporary buffers come from a fixed 128 MiB device array that each GPU holds for the life of
 * the process (about 350k pixels of ScatterBackup with one spatial neighbour; RS_E_SIZE_MISMATCH
 * beyond it).  Calls on one GPU share that array: calls that overlap on one GPU race. */
int32_t rs_render_frame_noalloc(rs_ctx* ctx, const rs_camera* cam, const rs_frame_io* io, void* stream);

/* The dictionary flags of the last frame (RS/ReservoirSplatting.cpp:503-506). */
int32_t rs_get_frame_flags(const rs_ctx* ctx, rs_frame_flags* out);

/* Scene::setFreeze (Source/Falcor/Scene/Scene.cpp:2518-2521) as the pass sees it: while
 * frozen, endFrame neither advances frameCount nor swaps the history
 * (RS/ReservoirSplatting.cpp:1672-1679), spatial reuse ping-pongs through temporary
 * buffers so the history survives (:2003-2004), and reset() is ignored (:813-820).  Every
 * frozen frame therefore re-renders the same frame from the same history with a new seed:
 * the "Replay for Accumulation" reference. */
int32_t rs_set_frozen(rs_ctx* ctx, int frozen);

/* reset(): frameCount = seed = 0; the next frame skips temporal reuse.  Ignored while frozen. */
int32_t rs_reset(rs_ctx* ctx);

/* Checkpoint of everything the next frame reads from earlier frames: the history
 * reservoirs and reconnection records, the camera keyframe ring, frameCount, seed and the
 * frame size.  rs_save_state with host = NULL returns the size in *bytes; otherwise *bytes
 * is the capacity on input and the size written on output.  rs_load_state restores a
 * checkpoint into a context with the same scene and options; the next frame then renders
 * bit for bit as it would have in the saved context. */
int32_t rs_save_state(rs_ctx* ctx, void* host, size_t* bytes);
int32_t rs_load_state(rs_ctx* ctx, const void* host, size_t bytes);

/* Per-stage device timing with hip events (off by default).  Timed frames run their stages in
 * order on the caller's stream, so that each pair of events brackets one stage. */
int32_t rs_set_timing(rs_ctx* ctx, int enable);
int32_t rs_get_stats(rs_ctx* ctx, rs_stats* out);

/* Synchronous device -> host copy of an internal buffer (RS_BUF_*); *bytes is the
 * capacity on input and the size written on output. */
int32_t rs_read_buffer(rs_ctx* ctx, int which, void* host, size_t* bytes);

/* Device pointer of an internal buffer (RS_BUF_*), valid until the next call that
 * resizes the context. */
int32_t rs_buffer_ptr(rs_ctx* ctx, int which, void** dptr, size_t* bytes);

/* VisualizeForwardReprojection (RS/VisualizeForwardReprojection.cs.slang:49-74, dump at
 * RS/ReservoirSplatting.cpp:468-493): for every pixel, the number of previous-frame samples
 * the last frame splatted into it (MultiScatter: the mean over its time partitions).
 * Writes width*height floats to `host`; *count is the capacity on input and the number
 * written on output.  All zeros when the last frame ran no forward reprojection. */
int32_t rs_scatter_counts(rs_ctx* ctx, float* host, size_t* count);

/* Frozen-reservoir replay and its adjoint (no counterpart in the reference; gradients of an
 * image loss with respect to material base colour and roughness).  rs_freeze snapshots the
 * last frame: each pixel's final reservoir (selected path and weight) and image value.
 * The replay R(theta) keeps the selection, weights, pdfs and the radiance arriving at the
 * reconnection vertex, and re-evaluates the BSDF factors at the primary hit and the
 * reconnection vertex with parameters theta (4 floats per material: base colour RGB,
 * roughness): R_c = colour_c * prod f_c(theta) / f_c(theta0).  At the scene's parameters
 * it reproduces the frozen image.  Device pointers with their element counts, which must be
 * numMaterials*4 for the parameters and gradients and width*height*4 for the images of the
 * frozen frame (RS_E_SIZE_MISMATCH otherwise); the output image has alpha 1;
 * rs_replay_backward writes dL/dtheta (doubles) for a given dL/dR image.  A new scene,
 * vertex or instance update drops the snapshot (RS_E_NO_SCENE until the next rs_freeze). */
int32_t rs_freeze(rs_ctx* ctx);
int32_t rs_replay_forward(rs_ctx* ctx, const float* d_params, size_t numParams, float* d_image, size_t imageFloats,
                          void* stream);
int32_t rs_replay_backward(rs_ctx* ctx, const float* d_params, size_t numParams, const float* d_grad_image,
                           size_t imageFloats, double* d_grad_params, void* stream);

/* Batched frames (no counterpart in the reference, which renders one scene per pass):
 * many independent contexts rendered per call as one hipGraph launch.  Every context must
 * have a scene; results are identical to calling rs_render_frame for each context.
 *
 * Fused form (the default, queues 0): the contexts are split into `groups` groups of about
 * equal pixel counts; each group renders with one launch per stage for all its contexts,
 * on a stream of its own, so groups overlap.  The launches are captured once per topology
 * (frame sizes, options, groups, tile width); a frame then costs the host one table upload
 * per group and one graph launch.  Stream form (queues 1..8): each context's frame is
 * recorded on one of `queues` side streams forked from `stream` (largest frame first onto
 * the least loaded), and the graph is re-recorded every frame and updated in place.
 *
 * rs_batch_configure sets the stream form's queue count (0 selects the fused form) and the
 * pixel-tile width of the per-pixel kernels (8, 16, 32 or 64; tiles are 256 pixels);
 * rs_batch_configure_fused selects the fused form with `groups` groups (1..8);
 * rs_batch_get_config reports the configuration (queues 0 in the fused form);
 * rs_batch_autotune renders `frames` frames per candidate (fused with 1, 2, 4, 6 or 8 groups,
 * streams with 1, 2, 4 or 8 queues, each at every tile width) and keeps the fastest. */
typedef struct rs_batch rs_batch;
int32_t rs_batch_create(rs_ctx* const* ctxs, uint32_t n, rs_batch** out);
void rs_batch_destroy(rs_batch* b);
const char* rs_batch_last_error(const rs_batch* b);
int32_t rs_batch_configure(rs_batch* b, uint32_t queues, uint32_t tileWidth);
int32_t rs_batch_configure_fused(rs_batch* b, uint32_t groups, uint32_t tileWidth);
int32_t rs_batch_get_config(const rs_batch* b, uint32_t* queues, uint32_t* groups, uint32_t* tileWidth);
int32_t rs_batch_render(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios, void* stream);
int32_t rs_batch_autotune(rs_batch* b, const rs_camera* cams, const rs_frame_io* ios, uint32_t frames, void* stream,
                          uint32_t* bestQueues, uint32_t* bestTileWidth, float* bestMs);

/* Temporal accumulation, the step after the pass: AccumulatePass
 * (Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:130-290, kernels
 * Accumulate.cs.slang accumulateSingle / accumulateSingleCompensated / accumulateDouble;
 * properties :57-94, defaults AccumulatePass.h:140-148).  The input and output are RGBA32F
 * device images of width*height pixels.  An accumulator owns its running sums; it resets
 * on rs_accum_reset, on a change of frame size and on any option change (the reference
 * resets on mode changes, :330-352).  Scene-change auto-reset (the reference's refresh
 * flags, :133-160) is the caller's: call rs_accum_reset.  execute():
 *   - with maxFrameCount > 0 and the count at it: overflow Stop returns without writing
 *     the output, Reset restarts, EMA keeps blending with weight 1 / (maxFrameCount + 1);
 *   - disabled: the input is copied to the output;
 *   - shouldAccumulate = 0 writes the input (or, with viewAccumulation, the current
 *     average) without adding the frame. */
#define RS_ACCUM_SINGLE 0
#define RS_ACCUM_SINGLE_COMPENSATED 1
#define RS_ACCUM_DOUBLE 2
#define RS_ACCUM_OVERFLOW_STOP 0
#define RS_ACCUM_OVERFLOW_RESET 1
#define RS_ACCUM_OVERFLOW_EMA 2
typedef struct rs_accum_options {
  uint32_t enabled;       /* default 1 */
  uint32_t precision;     /* RS_ACCUM_*, default Single */
  uint32_t maxFrameCount; /* 0 = no limit (default) */
  uint32_t overflowMode;  /* RS_ACCUM_OVERFLOW_*, default Stop */
} rs_accum_options;
typedef struct rs_accum rs_accum;
void rs_accum_default_options(rs_accum_options* out);
int32_t rs_accum_create(const rs_accum_options* opts, rs_accum** out);
void rs_accum_destroy(rs_accum* a);
const char* rs_accum_last_error(const rs_accum* a);
int32_t rs_accum_set_options(rs_accum* a, const rs_accum_options* opts);
int32_t rs_accum_reset(rs_accum* a);
uint32_t rs_accum_frame_count(const rs_accum* a);
int32_t rs_accum_execute(rs_accum* a, const float* d_src, float* d_dst, uint32_t width, uint32_t height,
                         int shouldAccumulate, int viewAccumulation, void* stream);

/* EmissivePowerSampler::generateAliasTable (Source/Falcor/Rendering/Lights/
 * EmissivePowerSampler.cpp:86-177) without the random shuffle: a deterministic Vose table
 * over `weights`, packed as two 32-bit words per slot (16-bit half threshold, two 24-bit
 * indices).  Host-only; rs_set_scene uses it for the light collection. */
int32_t rs_build_alias_table(const float* weights, uint32_t n, uint32_t* table, float* invWeightsSum);

#ifdef __cplusplus
}
#endif
#endif
