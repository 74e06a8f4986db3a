"""`ReservoirSplatting`: the render pass's property surface over the C ABI (include/rs.h)."""
import ctypes as C
import os
import warnings

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class RSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def library_path():
    """The in-tree HIP library; RS_AMD_LIBRARY selects another build of the same ABI."""
    return os.environ.get("RS_AMD_LIBRARY") or os.path.join(os.path.dirname(_HERE), "lib", "librs_amd.so")


def load_library():
    """Load the HIP library; raises if it has not been built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RSError(abi.RS_E_UNSUPPORTED, f"HIP library missing: {path} (run __graft_entry__.build())")
        # PyTorch ships its own HIP runtime.  Loaded after it, this library binds its HIP
        # symbols to that runtime, so the process has one; loaded first, it binds the system
        # runtime and PyTorch's own then finds no GPU.  Import torch first when it exists.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _LIB = abi.bind_product(C.CDLL(path))
    return _LIB


# Reference property keys (ReservoirSplatting.cpp:99-131) that this build accepts but
# that have no effect on the modelled path (texture LOD, alpha test, nested dielectrics,
# sample-generator choice, output sizing, colour format, light BVH).
_INERT_KEYS = {
    "sampleGenerator", "emissiveSampler", "lightBVHOptions", "useAlphaTest", "maxNestedMaterials",
    "useLightsInDielectricVolumes", "disableCaustics", "primaryLodMode", "lodBias", "outputSize",
    "fixedOutputSize", "colorFormat",
}


def _enum(value, table, name):
    if isinstance(value, str):
        if value not in table:
            raise RSError(abi.RS_E_INVALID_ARG, f"unknown {name} '{value}'")
        return table[value]
    return int(value)


class ReservoirSplatting:
    """createPass("ReservoirSplatting", props) over the MI355X library."""

    def __init__(self, props=None):
        self._lib = load_library()
        self._opts = abi.Options()
        self._lib.rs_default_options(C.byref(self._opts))
        self.warnings = []
        self._ctx = C.c_void_p()
        code = self._lib.rs_create(C.byref(self._opts), C.byref(self._ctx))
        if code not in (abi.RS_OK, abi.RS_E_BAD_OPTION):
            raise RSError(code, "rs_create failed")
        self._scene = None
        if props:
            self.setProperties(props)

    # -------------------------------------------------------------- lifetime
    def close(self):
        if self._ctx:
            self._lib.rs_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code):
        if code == abi.RS_E_BAD_OPTION:
            msg = self._lib.rs_last_error(self._ctx).decode()
            self.warnings.append(msg)
            warnings.warn(msg)
            return code
        if code != abi.RS_OK:
            raise RSError(code, self._lib.rs_last_error(self._ctx).decode())
        return code

    # -------------------------------------------------------------- properties
    def setProperties(self, props):
        """parseProperties + validateOptions (ReservoirSplatting.cpp:192-301)."""
        o = abi.Options()
        C.memmove(C.byref(o), C.byref(self._opts), C.sizeof(o))
        simple = {
            "samplesPerPixel": "samplesPerPixel", "maxSurfaceBounces": "maxSurfaceBounces",
            "maxDiffuseBounces": "maxDiffuseBounces", "maxSpecularBounces": "maxSpecularBounces",
            "maxTransmissionBounces": "maxTransmissionBounces", "useRussianRoulette": "useRussianRoulette",
            "misPowerExponent": "misPowerExponent", "adjustShadingNormals": "adjustShadingNormals",
            "specularRoughnessThreshold": "specularRoughnessThreshold",
            "enableTemporalResampling": "enableTemporalResampling",
            "enableSpatialResampling": "enableSpatialResampling", "numTimePartitions": "numTimePartitions",
            # keys beyond the reference's property list (its UI exposes them)
            "numSpatialIterations": "numSpatialIterations", "neighborCount": "neighborCount",
            "spatialGatherRadius": "spatialGatherRadius", "artificialFrameTime": "artificialFrameTime",
            "useConfidenceWeightsTemporally": "useConfidenceWeightsTemporally",
            "useConfidenceWeightsSpatially": "useConfidenceWeightsSpatially",
            "gbufferAdjustShadingNormals": "gbufferAdjustShadingNormals", "captureStages": "captureStages",
            "quirkMask": "quirkMask",
        }
        for key, value in props.items():
            if key in simple:
                setattr(o, simple[key], value)
            elif key == "fixedSeed":
                o.fixedSeed = int(value)
                o.useFixedSeed = 1
            elif key == "misHeuristic":
                o.misHeuristic = _enum(value, abi.MIS, key)
            elif key == "temporalReuse":
                o.temporalReuse = _enum(value, abi.TEMPORAL, key)
            elif key == "scatterBackupMISOption":
                o.scatterBackupMIS = _enum(value, abi.BACKUP_MIS, key)
            elif key == "gatherOption":
                o.gatherMechanism = _enum(value, abi.GATHER, key)
            elif key in _INERT_KEYS:
                pass
            else:
                msg = f"Unknown property '{key}' in ReservoirSplatting properties."
                self.warnings.append(msg)
                warnings.warn(msg)
        if "maxSurfaceBounces" in props:
            for k in ("maxDiffuseBounces", "maxSpecularBounces", "maxTransmissionBounces"):
                if k not in props:
                    setattr(o, k, o.maxSurfaceBounces)
        else:
            o.maxSurfaceBounces = max(o.maxDiffuseBounces, o.maxSpecularBounces, o.maxTransmissionBounces)
        code = self._lib.rs_set_options(self._ctx, C.byref(o))
        self._check(code)
        self._lib.rs_get_options(self._ctx, C.byref(self._opts))

    def getProperties(self):
        o = self._opts
        inv = lambda t, v: {b: a for a, b in t.items()}[v]
        return {
            "samplesPerPixel": o.samplesPerPixel, "maxSurfaceBounces": o.maxSurfaceBounces,
            "maxDiffuseBounces": o.maxDiffuseBounces, "maxSpecularBounces": o.maxSpecularBounces,
            "maxTransmissionBounces": o.maxTransmissionBounces, "useRussianRoulette": bool(o.useRussianRoulette),
            "misHeuristic": inv(abi.MIS, o.misHeuristic), "misPowerExponent": o.misPowerExponent,
            "adjustShadingNormals": bool(o.adjustShadingNormals),
            "specularRoughnessThreshold": o.specularRoughnessThreshold,
            "enableTemporalResampling": bool(o.enableTemporalResampling),
            "enableSpatialResampling": bool(o.enableSpatialResampling),
            "temporalReuse": inv(abi.TEMPORAL, o.temporalReuse),
            "scatterBackupMISOption": inv(abi.BACKUP_MIS, o.scatterBackupMIS),
            "gatherOption": inv(abi.GATHER, o.gatherMechanism), "numTimePartitions": o.numTimePartitions,
            "numSpatialIterations": o.numSpatialIterations, "neighborCount": o.neighborCount,
            "spatialGatherRadius": o.spatialGatherRadius, "artificialFrameTime": o.artificialFrameTime,
            **({"fixedSeed": o.fixedSeed} if o.useFixedSeed else {}),
            **({"quirkMask": o.quirkMask} if o.quirkMask != abi.QUIRKS_REFERENCE else {}),
        }

    @property
    def useFixedSeed(self):
        return bool(self._opts.useFixedSeed)

    @useFixedSeed.setter
    def useFixedSeed(self, v):
        o = abi.Options()
        C.memmove(C.byref(o), C.byref(self._opts), C.sizeof(o))
        o.useFixedSeed = int(bool(v))
        self._check(self._lib.rs_set_options(self._ctx, C.byref(o)))
        self._lib.rs_get_options(self._ctx, C.byref(self._opts))

    @property
    def fixedSeed(self):
        return int(self._opts.fixedSeed)

    @fixedSeed.setter
    def fixedSeed(self, v):
        o = abi.Options()
        C.memmove(C.byref(o), C.byref(self._opts), C.sizeof(o))
        o.fixedSeed = int(v)
        self._check(self._lib.rs_set_options(self._ctx, C.byref(o)))
        self._lib.rs_get_options(self._ctx, C.byref(self._opts))

    # -------------------------------------------------------------- scene / frames
    def setScene(self, scene):
        self._scene = scene
        self._check(self._lib.rs_set_scene(self._ctx, C.byref(scene.desc())))

    def setEnvMap(self, texels, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0)):
        """The scene's environment map: an (H, W, 3|4) float lat-long image, scaled by
        intensity * tint and rotated by Euler angles XYZ in degrees (Falcor EnvMap)."""
        d, keep = abi.env_map_desc(texels, intensity, tint, rotation)
        self._check(self._lib.rs_set_env_map(self._ctx, C.byref(d)))
        del keep

    def clearEnvMap(self):
        self._check(self._lib.rs_clear_env_map(self._ctx))

    def updateVertices(self, positions, normals):
        positions = np.ascontiguousarray(positions, np.float32)
        normals = np.ascontiguousarray(normals, np.float32)
        self._check(self._lib.rs_update_vertices(self._ctx, positions.ctypes.data, normals.ctypes.data,
                                                 positions.shape[0]))

    def rebuildBvh(self):
        """Full BVH build over the current vertices (updateVertices only refits)."""
        self._check(self._lib.rs_rebuild_bvh(self._ctx))

    def updateInstances(self, transforms):
        """Rigid instance motion: one row-major 3x4 transform per scene instance, applied on
        the GPU to the instance's object-space vertices (rs_update_instances)."""
        X = np.ascontiguousarray(transforms, np.float32).reshape(-1, 12)
        self._check(self._lib.rs_update_instances(self._ctx, X.ctypes.data, X.shape[0]))

    @property
    def frozen(self):
        """Scene::isFrozen as the pass sees it (rs_set_frozen)."""
        return bool(getattr(self, "_frozen", False))

    @frozen.setter
    def frozen(self, v):
        self._check(self._lib.rs_set_frozen(self._ctx, int(bool(v))))
        self._frozen = bool(v)

    def frameFlags(self):
        """The dictionary entries of the last frame (RS/ReservoirSplatting.cpp:503-506)."""
        f = abi.FrameFlags()
        self._check(self._lib.rs_get_frame_flags(self._ctx, C.byref(f)))
        return {"_suppressAutoReset": bool(f.suppressAutoReset), "_shouldAccumulate": bool(f.shouldAccumulate),
                "_viewAccumulation": bool(f.viewAccumulation)}

    def saveState(self):
        """Checkpoint of the history the next frame reads (rs_save_state), as bytes."""
        n = C.c_size_t(0)
        self._check(self._lib.rs_save_state(self._ctx, None, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        self._check(self._lib.rs_save_state(self._ctx, buf.ctypes.data, C.byref(n)))
        return buf.tobytes()

    def loadState(self, blob):
        buf = np.frombuffer(blob, np.uint8)
        self._check(self._lib.rs_load_state(self._ctx, buf.ctypes.data, buf.size))

    def reset(self):
        self._check(self._lib.rs_reset(self._ctx))

    def execute(self, camera, width, height, mvec=None, color=None, stream=None, vbuffer=None,
                vbufferPrimitiveIndexBits=0, rayCount=None, pathLength=None, sceneFlags=0):
        """One frame.  The buffers are device pointers (ints) or torch CUDA tensors, which are
        checked for size, dtype and contiguity: mvec (H, W, 2) float32, color (H, W, 4) float32,
        vbuffer (H, W, 4) int32/uint32 PackedHitInfo, rayCount and pathLength (H, W) int32/uint32."""
        P = int(width) * int(height)
        io = abi.FrameIO()
        io.width, io.height = int(width), int(height)
        io.d_mvec = _dptr(mvec, 2 * P, _F32, "mvec")
        io.d_color = _dptr(color, 4 * P, _F32, "color")
        io.d_vbuffer = _dptr(vbuffer, 4 * P, _U32, "vbuffer")
        io.vbufferPrimitiveIndexBits = int(vbufferPrimitiveIndexBits)
        io.d_rayCount = _dptr(rayCount, P, _U32, "rayCount")
        io.d_pathLength = _dptr(pathLength, P, _U32, "pathLength")
        io.sceneFlags = int(sceneFlags)
        st = None
        if stream is not None:
            st = C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        return self._check(self._lib.rs_render_frame(self._ctx, C.byref(camera), C.byref(io), st))

    def set_timing(self, enable=True):
        self._check(self._lib.rs_set_timing(self._ctx, int(bool(enable))))

    def stats(self):
        s = abi.Stats()
        self._check(self._lib.rs_get_stats(self._ctx, C.byref(s)))
        return s

    @property
    def pixelStats(self):
        s = self.stats()
        return {"closestHitRays": int(s.raysClosestHit), "visibilityRays": int(s.raysVisibility),
                "totalRays": int(s.raysClosestHit + s.raysVisibility)}

    def read_buffer(self, name, width, height, partitions=1):
        """Raw bytes of an internal buffer; the splat lists hold `partitions` x W x H entries."""
        which = abi.BUF[name] if isinstance(name, str) else int(name)
        n = abi.BUF_BYTES_PER_PIXEL[which] * width * height * (partitions if which in abi.SPLAT_BUFS else 1)
        out = np.empty(n, np.uint8)
        size = C.c_size_t(n)
        self._check(self._lib.rs_read_buffer(self._ctx, which, out.ctypes.data, C.byref(size)))
        return out

    # -------------------------------------------------------------- frozen replay / adjoint
    def freeze(self):
        """Snapshot the last frame's final reservoirs for the replay and its adjoint."""
        self._check(self._lib.rs_freeze(self._ctx))

    def materialParams(self):
        """The scene's (base colour RGB, roughness) per material as a float32 array."""
        return np.asarray([list(m[0]) + [m[1]] for m in self._scene.materials], np.float32)

    def replay_forward(self, params, image, stream=None):
        """R(params) into `image` (device tensors: params (numMaterials, 4) float32, image
        (H, W, 4) float32 of the frozen frame; the library checks both sizes)."""
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        self._check(self._lib.rs_replay_forward(self._ctx, _dptr(params, None, _F32, "params"), _numel(params),
                                                _dptr(image, None, _F32, "image"), _numel(image), st))

    def replay_backward(self, params, grad_image, grad_params, stream=None):
        """dL/dparams (float64 device tensor (numMaterials, 4)) for dL/dR = grad_image."""
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        if _numel(grad_params) != _numel(params):
            raise RSError(abi.RS_E_SIZE_MISMATCH, "grad_params must have the shape of params")
        self._check(self._lib.rs_replay_backward(self._ctx, _dptr(params, None, _F32, "params"), _numel(params),
                                                 _dptr(grad_image, None, _F32, "grad_image"), _numel(grad_image),
                                                 _dptr(grad_params, None, _F64, "grad_params"), st))

    def scatterCounts(self, width, height):
        """Splats received per pixel by the last frame (VisualizeForwardReprojection)."""
        out = np.empty(width * height, np.float32)
        n = C.c_size_t(out.size)
        self._check(self._lib.rs_scatter_counts(self._ctx, out.ctypes.data, C.byref(n)))
        return out.reshape(height, width)

    def dumpScatterCount(self, path, width, height):
        """The reference's scatter-count dump (ReservoirSplatting.cpp:468-493): one value per
        line, in ostream's default float format."""
        with open(path, "w") as f:
            for v in self.scatterCounts(width, height).ravel():
                f.write(f"{float(v):.6g}\n")

    def buffer_ptr(self, name):
        which = abi.BUF[name] if isinstance(name, str) else int(name)
        ptr, size = C.c_void_p(), C.c_size_t()
        self._check(self._lib.rs_buffer_ptr(self._ctx, which, C.byref(ptr), C.byref(size)))
        return ptr.value, size.value


_F32 = ("float32",)
_F64 = ("float64",)
_U32 = ("int32", "uint32")


def _numel(x):
    return int(x.numel()) if hasattr(x, "numel") else 0


def _dptr(x, numel=None, dtypes=None, name="buffer"):
    """Device address of a buffer: an int is taken as is; a torch tensor must be a contiguous
    CUDA tensor of one of `dtypes` with `numel` elements (when given)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_cuda:
            raise RSError(abi.RS_E_INVALID_ARG, f"{name}: device tensors only")
        if not x.is_contiguous():
            raise RSError(abi.RS_E_INVALID_ARG, f"{name}: the tensor must be contiguous")
        if dtypes is not None and str(x.dtype).replace("torch.", "") not in dtypes:
            raise RSError(abi.RS_E_INVALID_ARG, f"{name}: dtype {x.dtype}, expected one of {dtypes}")
        if numel is not None and x.numel() != numel:
            raise RSError(abi.RS_E_SIZE_MISMATCH, f"{name}: {x.numel()} elements, expected {numel}")
        return x.data_ptr()
    raise RSError(abi.RS_E_INVALID_ARG, f"unsupported buffer type {type(x)}")
