"""ctypes mirror of include/rs.h (structs, constants, function signatures).

The structs are field-for-field copies of the C declarations; `bind(lib, prefix)`
attaches argtypes/restype for the `rs_*` product entry points.
"""
import ctypes as C

RS_OK = 0
RS_E_INVALID_ARG = -1
RS_E_FRAME_TOO_LARGE = -2
RS_E_SIZE_MISMATCH = -3
RS_E_NO_SCENE = -4
RS_E_BAD_OPTION = -5
RS_E_HIP = -6
RS_E_OUT_OF_MEMORY = -7
RS_E_UNSUPPORTED = -8

ERROR_NAMES = {
    0: "RS_OK", -1: "RS_E_INVALID_ARG", -2: "RS_E_FRAME_TOO_LARGE", -3: "RS_E_SIZE_MISMATCH",
    -4: "RS_E_NO_SCENE", -5: "RS_E_BAD_OPTION", -6: "RS_E_HIP", -7: "RS_E_OUT_OF_MEMORY",
    -8: "RS_E_UNSUPPORTED",
}

TEMPORAL = {"GatherOnly": 0, "ScatterOnly": 1, "ScatterBackup": 2, "MultiScatter": 3}
GATHER = {"Fast": 0, "Clamped": 1, "Robust": 2}
BACKUP_MIS = {"Balance": 0, "Pairwise": 1}
MIS = {"Balance": 0, "PowerTwo": 1, "PowerExp": 2}

DIFFUSE_FROSTBITE = 0
DIFFUSE_LAMBERT = 1

LOBE_DIFFUSE_REFLECTION = 0x01
LOBE_SPECULAR_REFLECTION = 0x02
LOBE_ALL = 0xFF

BUF = {
    "history_reservoirs": 0, "history_recon": 1, "temporal_reservoirs": 2, "temporal_recon": 3,
    "initial_reservoirs": 4, "initial_recon": 5, "backup_reservoirs": 6, "backup_recon": 7,
    "mvec": 8, "vbuffer": 9, "floating_coords": 10, "splat_dst": 11, "cell_count": 12,
    "cell_offset": 13, "sorted": 14, "color": 15, "final_reservoirs": 16, "final_recon": 17,
}
BUF_BYTES_PER_PIXEL = {0: 32, 1: 112, 2: 32, 3: 112, 4: 32, 5: 112, 6: 32, 7: 112, 8: 8, 9: 12,
                       10: 8, 11: 4, 12: 4, 13: 4, 14: 4, 15: 16, 16: 32, 17: 112}

# rs_frame_io.sceneFlags (what the pass reads from the Scene each frame)
SCENE_SHOULD_RESET = 0x1
SCENE_REPLAYING = 0x2
SCENE_ACCUMULATION_FRAME = 0x4
SCENE_VIEW_ACCUMULATION = 0x8

SPLAT_BUFS = (11, 12, 13, 14)  # splat_dst, cell_count, cell_offset, sorted: K partitions x P

STAGES = ["vbuffer", "initial", "collect", "reproject", "sort", "temporal", "spatial", "resolve"]


class Options(C.Structure):
    _fields_ = [
        ("samplesPerPixel", C.c_uint32), ("maxSurfaceBounces", C.c_uint32),
        ("maxDiffuseBounces", C.c_uint32), ("maxSpecularBounces", C.c_uint32),
        ("maxTransmissionBounces", C.c_uint32), ("useRussianRoulette", C.c_uint32),
        ("misHeuristic", C.c_uint32), ("misPowerExponent", C.c_float),
        ("specularRoughnessThreshold", C.c_float), ("adjustShadingNormals", C.c_uint32),
        ("gbufferAdjustShadingNormals", C.c_uint32), ("enableTemporalResampling", C.c_uint32),
        ("temporalReuse", C.c_uint32), ("gatherMechanism", C.c_uint32),
        ("scatterBackupMIS", C.c_uint32), ("numTimePartitions", C.c_uint32),
        ("useConfidenceWeightsTemporally", C.c_uint32), ("enableSpatialResampling", C.c_uint32),
        ("numSpatialIterations", C.c_uint32), ("neighborCount", C.c_uint32),
        ("useConfidenceWeightsSpatially", C.c_uint32), ("spatialGatherRadius", C.c_float),
        ("artificialFrameTime", C.c_float), ("useFixedSeed", C.c_uint32), ("fixedSeed", C.c_uint32),
        ("captureStages", C.c_uint32), ("quirkMask", C.c_uint32),
    ]


class Material(C.Structure):
    _fields_ = [("baseColor", C.c_float * 3), ("roughness", C.c_float), ("emissive", C.c_float * 3),
                ("metallic", C.c_float), ("ior", C.c_float), ("activeLobes", C.c_uint32),
                ("doubleSided", C.c_uint32), ("_pad", C.c_uint32)]


class InstanceDesc(C.Structure):
    _fields_ = [("firstVertex", C.c_uint32), ("numVertices", C.c_uint32), ("firstTriangle", C.c_uint32),
                ("numTriangles", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("numVertices", C.c_uint32),
                ("indices", C.c_void_p), ("triMaterial", C.c_void_p), ("numTriangles", C.c_uint32),
                ("materials", C.c_void_p), ("numMaterials", C.c_uint32), ("diffuseModel", C.c_uint32),
                ("instances", C.c_void_p), ("numInstances", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("focalLength", C.c_float),
                ("frameHeight", C.c_float), ("aspectRatio", C.c_float), ("focalDistance", C.c_float),
                ("apertureRadius", C.c_float), ("shutterSpeed", C.c_float)]


class FrameIO(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("d_mvec", C.c_void_p),
                ("d_color", C.c_void_p), ("d_vbuffer", C.c_void_p), ("vbufferPrimitiveIndexBits", C.c_uint32),
                ("sceneFlags", C.c_uint32), ("d_rayCount", C.c_void_p), ("d_pathLength", C.c_void_p)]


class FrameFlags(C.Structure):
    _fields_ = [("suppressAutoReset", C.c_uint32), ("shouldAccumulate", C.c_uint32),
                ("viewAccumulation", C.c_uint32)]


class CpuFrameExtra(C.Structure):
    """The oracle's host-memory counterpart of rs_frame_io's optional fields."""
    _fields_ = [("vbuffer", C.c_void_p), ("vbufferPrimitiveIndexBits", C.c_uint32), ("sceneFlags", C.c_uint32),
                ("rayCount", C.c_void_p), ("pathLength", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("raysClosestHit", C.c_uint64), ("raysVisibility", C.c_uint64),
                ("frameCount", C.c_uint32), ("seed", C.c_uint32), ("msTotal", C.c_float),
                ("msStage", C.c_float * 12), ("bvhNodes", C.c_uint32), ("bvhLevels", C.c_uint32),
                ("bvhBuild", C.c_uint32)]


BVH_BUILD_CLUSTERING = 0
BVH_BUILD_BALANCED = 1


class EnvMapDesc(C.Structure):
    _fields_ = [("texels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("intensity", C.c_float),
                ("tint", C.c_float * 3), ("rotation", C.c_float * 3)]


class AccumOptions(C.Structure):
    _fields_ = [("enabled", C.c_uint32), ("precision", C.c_uint32), ("maxFrameCount", C.c_uint32),
                ("overflowMode", C.c_uint32)]


# Reference quirks reproduced while set in Options.quirkMask (include/rs.h RS_QUIRK_*).
QUIRK = {"robustGrey": 0x1, "pathTimeZero": 0x2, "multiProjection": 0x4, "misDropsLightHits": 0x8,
         "sppLastSample": 0x10}
QUIRKS_REFERENCE = 0x1F

ACCUM_PRECISION = {"Single": 0, "SingleCompensated": 1, "Double": 2}
ACCUM_OVERFLOW = {"Stop": 0, "Reset": 1, "EMA": 2}


# Every symbol include/rs.h declares (checked by the CPU test suite).
PRODUCT_SYMBOLS = [
    "rs_default_options", "rs_create", "rs_destroy", "rs_last_error", "rs_set_options",
    "rs_get_options", "rs_set_scene", "rs_update_vertices", "rs_render_frame", "rs_reset",
    "rs_set_timing", "rs_get_stats", "rs_read_buffer", "rs_buffer_ptr", "rs_build_alias_table",
    "rs_scatter_counts", "rs_rebuild_bvh", "rs_freeze", "rs_replay_forward", "rs_replay_backward",
    "rs_batch_create", "rs_batch_destroy", "rs_batch_last_error", "rs_batch_configure", "rs_batch_configure_fused",
    "rs_batch_get_config", "rs_batch_render",
    "rs_batch_autotune", "rs_accum_default_options", "rs_accum_create", "rs_accum_destroy", "rs_accum_last_error",
    "rs_accum_set_options", "rs_accum_reset", "rs_accum_frame_count", "rs_accum_execute",
    "rs_set_env_map", "rs_clear_env_map", "rs_update_instances", "rs_get_frame_flags", "rs_set_frozen",
    "rs_save_state", "rs_load_state", "rs_render_frame_noalloc",
]


def bind_product(lib):
    P, I32, U32 = C.c_void_p, C.c_int32, C.c_uint32
    sig = {
        "rs_default_options": (None, [C.POINTER(Options)]),
        "rs_create": (I32, [C.POINTER(Options), C.POINTER(P)]),
        "rs_destroy": (None, [P]),
        "rs_last_error": (C.c_char_p, [P]),
        "rs_set_options": (I32, [P, C.POINTER(Options)]),
        "rs_get_options": (I32, [P, C.POINTER(Options)]),
        "rs_set_scene": (I32, [P, C.POINTER(SceneDesc)]),
        "rs_update_vertices": (I32, [P, P, P, U32]),
        "rs_render_frame": (I32, [P, C.POINTER(Camera), C.POINTER(FrameIO), P]),
        "rs_render_frame_noalloc": (I32, [P, C.POINTER(Camera), C.POINTER(FrameIO), P]),
        "rs_reset": (I32, [P]),
        "rs_set_timing": (I32, [P, C.c_int]),
        "rs_get_stats": (I32, [P, C.POINTER(Stats)]),
        "rs_read_buffer": (I32, [P, C.c_int, P, C.POINTER(C.c_size_t)]),
        "rs_buffer_ptr": (I32, [P, C.c_int, C.POINTER(P), C.POINTER(C.c_size_t)]),
        "rs_build_alias_table": (I32, [P, U32, P, C.POINTER(C.c_float)]),
        "rs_scatter_counts": (I32, [P, P, C.POINTER(C.c_size_t)]),
        "rs_rebuild_bvh": (I32, [P]),
        "rs_freeze": (I32, [P]),
        "rs_replay_forward": (I32, [P, P, C.c_size_t, P, C.c_size_t, P]),
        "rs_replay_backward": (I32, [P, P, C.c_size_t, P, C.c_size_t, P, P]),
        "rs_batch_create": (I32, [P, U32, C.POINTER(P)]),
        "rs_batch_destroy": (None, [P]),
        "rs_batch_last_error": (C.c_char_p, [P]),
        "rs_batch_configure": (I32, [P, U32, U32]),
        "rs_batch_configure_fused": (I32, [P, U32, U32]),
        "rs_batch_get_config": (I32, [P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32)]),
        "rs_batch_render": (I32, [P, P, P, P]),
        "rs_batch_autotune": (I32, [P, P, P, U32, P, C.POINTER(U32), C.POINTER(U32), C.POINTER(C.c_float)]),
        "rs_accum_default_options": (None, [C.POINTER(AccumOptions)]),
        "rs_accum_create": (I32, [C.POINTER(AccumOptions), C.POINTER(P)]),
        "rs_accum_destroy": (None, [P]),
        "rs_accum_last_error": (C.c_char_p, [P]),
        "rs_accum_set_options": (I32, [P, C.POINTER(AccumOptions)]),
        "rs_accum_reset": (I32, [P]),
        "rs_accum_frame_count": (U32, [P]),
        "rs_accum_execute": (I32, [P, P, P, U32, U32, C.c_int, C.c_int, P]),
        "rs_set_env_map": (I32, [P, C.POINTER(EnvMapDesc)]),
        "rs_clear_env_map": (I32, [P]),
        "rs_update_instances": (I32, [P, P, U32]),
        "rs_get_frame_flags": (I32, [P, C.POINTER(FrameFlags)]),
        "rs_set_frozen": (I32, [P, C.c_int]),
        "rs_save_state": (I32, [P, P, C.POINTER(C.c_size_t)]),
        "rs_load_state": (I32, [P, P, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


class ImportCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("focalLength", C.c_float), ("frameHeight", C.c_float), ("apertureRadius", C.c_float),
                ("focalDistance", C.c_float), ("present", C.c_uint32)]


# Scene ingestion (include/rs_import.h).
RS_E_IMPORT = -9
PRIM = {"quad": 0, "disk": 1, "cube": 2, "sphere": 3}

IMPORT_SYMBOLS = [
    "rs_builder_create", "rs_builder_destroy", "rs_builder_log", "rs_builder_import", "rs_builder_add_material",
    "rs_builder_add_mesh", "rs_builder_add_primitive", "rs_builder_add_mesh_file", "rs_builder_add_instance",
    "rs_builder_set_camera", "rs_builder_set_env_map", "rs_builder_set_env_map_file", "rs_builder_set_film",
    "rs_builder_set_diffuse_model", "rs_builder_get_scene", "rs_builder_get_camera", "rs_builder_get_env_map",
    "rs_builder_get_film",
]


def bind_import(lib):
    P, I32, U32, F = C.c_void_p, C.c_int32, C.c_uint32, C.POINTER(C.c_float)
    sig = {
        "rs_builder_create": (P, []),
        "rs_builder_destroy": (None, [P]),
        "rs_builder_log": (C.c_char_p, [P]),
        "rs_builder_import": (I32, [P, C.c_char_p]),
        "rs_builder_add_material": (I32, [P, C.POINTER(Material)]),
        "rs_builder_add_mesh": (I32, [P, P, P, U32, P, U32, P, U32, U32]),
        "rs_builder_add_primitive": (I32, [P, U32, P, P, U32]),
        "rs_builder_add_mesh_file": (I32, [P, C.c_char_p, P, U32, U32]),
        "rs_builder_add_instance": (I32, [P, U32, P, U32]),
        "rs_builder_set_camera": (I32, [P, C.POINTER(ImportCamera)]),
        "rs_builder_set_env_map": (I32, [P, C.POINTER(EnvMapDesc)]),
        "rs_builder_set_env_map_file": (I32, [P, C.c_char_p, C.c_float, F, F, U32]),
        "rs_builder_set_film": (I32, [P, U32, U32]),
        "rs_builder_set_diffuse_model": (I32, [P, U32]),
        "rs_builder_get_scene": (I32, [P, C.POINTER(SceneDesc)]),
        "rs_builder_get_camera": (I32, [P, C.POINTER(ImportCamera)]),
        "rs_builder_get_env_map": (I32, [P, C.POINTER(EnvMapDesc)]),
        "rs_builder_get_film": (I32, [P, C.POINTER(U32), C.POINTER(U32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def bind_oracle(lib):
    P, I32, U32 = C.c_void_p, C.c_int32, C.c_uint32
    sig = {
        "rs_cpu_create": (I32, [C.POINTER(Options), C.POINTER(P)]),
        "rs_cpu_destroy": (None, [P]),
        "rs_cpu_set_threads": (I32, [C.c_int]),
        "rs_cpu_set_scene": (I32, [P, C.POINTER(SceneDesc)]),
        "rs_cpu_update_vertices": (I32, [P, P, P, U32]),
        "rs_cpu_reset": (I32, [P]),
        "rs_cpu_trace": (I32, [P, P, U32, P]),
        "rs_cpu_render_frame": (I32, [P, C.POINTER(Camera), U32, U32, P, P, C.POINTER(CpuFrameExtra)]),
        "rs_cpu_update_instances": (I32, [P, P, U32]),
        "rs_cpu_get_vertices": (I32, [P, P, P, U32]),
        "rs_cpu_set_frozen": (I32, [P, C.c_int]),
        "rs_cpu_get_frame_flags": (I32, [P, C.POINTER(FrameFlags)]),
        "rs_cpu_get_stats": (I32, [P, C.POINTER(Stats)]),
        "rs_cpu_read_buffer": (I32, [P, C.c_int, P, C.POINTER(C.c_size_t)]),
        "rs_cpu_scatter_counts": (I32, [P, P, C.POINTER(C.c_size_t)]),
        "rs_cpu_replay": (I32, [P, P, P, P, P, P]),
        "rs_cpu_accum_create": (I32, [C.POINTER(AccumOptions), C.POINTER(P)]),
        "rs_cpu_accum_destroy": (None, [P]),
        "rs_cpu_accum_set_options": (I32, [P, C.POINTER(AccumOptions)]),
        "rs_cpu_accum_reset": (I32, [P]),
        "rs_cpu_accum_frame_count": (U32, [P]),
        "rs_cpu_accum_execute": (I32, [P, P, P, U32, U32, C.c_int, C.c_int]),
        "rs_cpu_set_env_map": (I32, [P, C.POINTER(EnvMapDesc)]),
        "rs_cpu_clear_env_map": (I32, [P]),
        "rs_cpu_env_query": (I32, [P, P, U32, P]),
        "rs_cpu_env_sample": (I32, [P, P, U32, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def env_map_desc(texels, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0)):
    """EnvMapDesc over an (H, W, 3|4) float image (row 0 at the top); returns (desc, the
    contiguous RGBA32F array it points to, which the caller keeps alive during the call)."""
    import numpy as np
    t = np.asarray(texels, np.float32)
    if t.ndim != 3 or t.shape[2] not in (3, 4):
        raise ValueError("environment map must be (H, W, 3) or (H, W, 4)")
    if t.shape[2] == 3:
        t = np.concatenate([t, np.ones(t.shape[:2] + (1,), np.float32)], axis=2)
    t = np.ascontiguousarray(t)
    d = EnvMapDesc()
    d.texels = t.ctypes.data
    d.width, d.height = t.shape[1], t.shape[0]
    d.intensity = float(intensity)
    d.tint[:] = [float(x) for x in tint]
    d.rotation[:] = [float(x) for x in rotation]
    return d, t
