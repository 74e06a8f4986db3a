"""Test-side wrapper of the CPU oracle (oracle/_build/librs_oracle.so) and buffer decoding.

The oracle is the checker: this module is only imported by tests, smoke() and the
bench's cpu_baseline leg.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "reservoir-splatting_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from rs_amd import abi  # noqa: E402

ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "librs_oracle.so")
_LIB = None


def oracle_lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"oracle not built: {ORACLE_PATH}")
        _LIB = abi.bind_oracle(C.CDLL(ORACLE_PATH))
    return _LIB


class OraclePass:
    """Same calls as rs_amd.ReservoirSplatting, executed by the CPU oracle."""

    def __init__(self, opts):
        self._lib = oracle_lib()
        self._opts = abi.Options()
        C.memmove(C.byref(self._opts), C.byref(opts), C.sizeof(opts))
        self._ctx = C.c_void_p()
        assert self._lib.rs_cpu_create(C.byref(self._opts), C.byref(self._ctx)) == 0
        self.color = None

    def __del__(self):
        try:
            if self._ctx:
                self._lib.rs_cpu_destroy(self._ctx)
        except Exception:
            pass

    def setScene(self, scene):
        self._scene = scene
        assert self._lib.rs_cpu_set_scene(self._ctx, C.byref(scene.desc())) == 0

    def setEnvMap(self, texels, intensity=1.0, tint=(1.0, 1.0, 1.0), rotation=(0.0, 0.0, 0.0)):
        d, keep = abi.env_map_desc(texels, intensity, tint, rotation)
        assert self._lib.rs_cpu_set_env_map(self._ctx, C.byref(d)) == 0
        del keep

    def clearEnvMap(self):
        assert self._lib.rs_cpu_clear_env_map(self._ctx) == 0

    def envQuery(self, dirs):
        """(n, 3) directions -> (n, 4): radiance RGB and the sampling pdf."""
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.empty((d.shape[0], 4), np.float32)
        assert self._lib.rs_cpu_env_query(self._ctx, d.ctypes.data, d.shape[0], out.ctypes.data) == 0
        return out

    def envSample(self, rnd):
        """(n, 2) uniforms -> (n, 8): direction, pdf, radiance RGB, 0."""
        r = np.ascontiguousarray(rnd, np.float32)
        out = np.empty((r.shape[0], 8), np.float32)
        assert self._lib.rs_cpu_env_sample(self._ctx, r.ctypes.data, r.shape[0], out.ctypes.data) == 0
        return out

    def updateVertices(self, positions, normals):
        positions = np.ascontiguousarray(positions, np.float32)
        normals = np.ascontiguousarray(normals, np.float32)
        assert self._lib.rs_cpu_update_vertices(self._ctx, positions.ctypes.data, normals.ctypes.data,
                                                positions.shape[0]) == 0

    def updateInstances(self, transforms):
        X = np.ascontiguousarray(transforms, np.float32).reshape(-1, 12)
        assert self._lib.rs_cpu_update_instances(self._ctx, X.ctypes.data, X.shape[0]) == 0

    def vertices(self):
        """The current world-space (positions, normals), (numVertices, 3) float32 each."""
        n = self._scene.positions.shape[0]
        P = np.empty((n, 3), np.float32)
        N = np.empty((n, 3), np.float32)
        assert self._lib.rs_cpu_get_vertices(self._ctx, P.ctypes.data, N.ctypes.data, n) == 0
        return P, N

    @property
    def frozen(self):
        return bool(getattr(self, "_frozen", False))

    @frozen.setter
    def frozen(self, v):
        assert self._lib.rs_cpu_set_frozen(self._ctx, int(bool(v))) == 0
        self._frozen = bool(v)

    def frameFlags(self):
        f = abi.FrameFlags()
        assert self._lib.rs_cpu_get_frame_flags(self._ctx, C.byref(f)) == 0
        return {"_suppressAutoReset": bool(f.suppressAutoReset), "_shouldAccumulate": bool(f.shouldAccumulate),
                "_viewAccumulation": bool(f.viewAccumulation)}

    def reset(self):
        assert self._lib.rs_cpu_reset(self._ctx) == 0

    def execute(self, camera, width, height, mvec=None, vbuffer=None, vbufferPrimitiveIndexBits=0, stats=False,
                sceneFlags=0):
        """One frame; with stats=True also the per-pixel ray counts and path lengths
        (self.rayCount, self.pathLength, (H, W) uint32)."""
        self.color = np.zeros((height, width, 4), np.float32)
        mv = None if mvec is None else np.ascontiguousarray(mvec, np.float32).ctypes.data
        ex = abi.CpuFrameExtra()
        vb = None
        if vbuffer is not None:
            vb = np.ascontiguousarray(vbuffer, np.uint32)
            ex.vbuffer = vb.ctypes.data
        ex.vbufferPrimitiveIndexBits = int(vbufferPrimitiveIndexBits)
        ex.sceneFlags = int(sceneFlags)
        if stats:
            self.rayCount = np.zeros((height, width), np.uint32)
            self.pathLength = np.full((height, width), 7, np.uint32)
            ex.rayCount = self.rayCount.ctypes.data
            ex.pathLength = self.pathLength.ctypes.data
        code = self._lib.rs_cpu_render_frame(self._ctx, C.byref(camera), width, height, mv, self.color.ctypes.data,
                                             C.byref(ex))
        if code != 0:
            raise RuntimeError(f"oracle frame failed: {abi.ERROR_NAMES.get(code, code)}")
        return code

    def stats(self):
        s = abi.Stats()
        assert self._lib.rs_cpu_get_stats(self._ctx, C.byref(s)) == 0
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
        code = self._lib.rs_cpu_read_buffer(self._ctx, which, out.ctypes.data, C.byref(size))
        assert code == 0, code
        return out

    def replay(self, params, width, height, target=None):
        """FP64 frozen replay of the last frame: (image (H, W, 3), loss, grad (numMaterials, 4))."""
        params = np.ascontiguousarray(params, np.float64)
        img = np.empty((height, width, 3), np.float64)
        grad = np.zeros(params.shape, np.float64)
        loss = C.c_double(0.0)
        tgt = None if target is None else np.ascontiguousarray(target, np.float32)
        assert self._lib.rs_cpu_replay(self._ctx, params.ctypes.data, None if tgt is None else tgt.ctypes.data,
                                       img.ctypes.data, grad.ctypes.data, C.byref(loss)) == 0
        return img, loss.value, grad

    def scatterCounts(self, width, height):
        out = np.empty(width * height, np.float32)
        n = C.c_size_t(out.size)
        assert self._lib.rs_cpu_scatter_counts(self._ctx, out.ctypes.data, C.byref(n)) == 0
        return out.reshape(height, width)

    def trace(self, rays):
        """rays: (n, 8) float32 [ox, oy, oz, tmin, dx, dy, dz, tmax] -> (n, 3) uint32 [prim, u bits, v bits]."""
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.empty((rays.shape[0], 3), np.uint32)
        assert self._lib.rs_cpu_trace(self._ctx, rays.ctypes.data, rays.shape[0], out.ctypes.data) == 0
        return out


def set_threads(n=0):
    return oracle_lib().rs_cpu_set_threads(int(n))


RESERVOIR_DTYPE = np.dtype([("confidence", "<f4"), ("totalWeight", "<f4"), ("integrand", "<f4", 3),
                            ("subPixel", "<f4", 2), ("_pad", "<f4")])


def decode_reservoirs(raw):
    return np.frombuffer(raw.tobytes(), RESERVOIR_DTYPE)


def decode_recon(raw):
    w = np.frombuffer(raw.tobytes(), np.uint32).reshape(-1, 28)
    f = w.view(np.float32)
    return {
        "subPixel": f[:, 0:2], "lensSample": f[:, 2:4], "time": f[:, 4], "pathLength": w[:, 5] & 0xFF,
        "firstComp": (w[:, 5] >> 8) & 0xF, "secondComp": (w[:, 5] >> 12) & 0xF,
        "transmissionEvent": (w[:, 5] >> 16) & 1, "lightIsNEE": (w[:, 5] >> 17) & 1,
        "lightIsDistant": (w[:, 5] >> 18) & 1, "lightPdf": f[:, 6], "subPixelJacobian": f[:, 7],
        "firstHitPrim": w[:, 8], "firstHitUV": f[:, 9:11], "lensVertexJacobian": f[:, 11],
        "firstWi": f[:, 12:15], "secondaryPathJacobian": f[:, 15], "secondHitPrim": w[:, 16],
        "secondHitUV": f[:, 17:19], "irradiance": np.stack([f[:, 19], f[:, 23], f[:, 24]], 1),
        "secondWo": f[:, 20:23], "earlyThroughput": f[:, 25:28],
    }


def first_mismatch(a, b, width, bytes_per_pixel):
    """(pixel index, word index) of the first differing 32-bit word, or None."""
    wa = np.frombuffer(a.tobytes(), np.uint32).reshape(-1, bytes_per_pixel // 4)
    wb = np.frombuffer(b.tobytes(), np.uint32).reshape(-1, bytes_per_pixel // 4)
    d = np.argwhere(wa != wb)
    if d.size == 0:
        return None
    p, k = d[0]
    return int(p), int(k), (int(p) % width, int(p) // width), int(wa[p, k]), int(wb[p, k]), int(len(np.unique(d[:, 0])))


class OracleAccum:
    """rs_cpu_accum_*: the CPU restatement of AccumulatePass, on host RGBA32F images."""

    def __init__(self, enabled=1, precision=0, maxFrameCount=0, overflowMode=0):
        self._lib = oracle_lib()
        self._opts = abi.AccumOptions(enabled, precision, maxFrameCount, overflowMode)
        self._h = C.c_void_p()
        assert self._lib.rs_cpu_accum_create(C.byref(self._opts), C.byref(self._h)) == 0

    def __del__(self):
        try:
            if self._h:
                self._lib.rs_cpu_accum_destroy(self._h)
        except Exception:
            pass

    def set_options(self, enabled=1, precision=0, maxFrameCount=0, overflowMode=0):
        self._opts = abi.AccumOptions(enabled, precision, maxFrameCount, overflowMode)
        assert self._lib.rs_cpu_accum_set_options(self._h, C.byref(self._opts)) == 0

    def reset(self):
        assert self._lib.rs_cpu_accum_reset(self._h) == 0

    @property
    def frameCount(self):
        return int(self._lib.rs_cpu_accum_frame_count(self._h))

    def execute(self, src, dst, shouldAccumulate=True, viewAccumulation=False):
        """src, dst: C-contiguous float32 arrays of shape (H, W, 4); dst is written in place."""
        assert src.dtype == np.float32 and dst.dtype == np.float32 and src.flags.c_contiguous and dst.flags.c_contiguous
        H, W = src.shape[0], src.shape[1]
        assert self._lib.rs_cpu_accum_execute(self._h, src.ctypes.data, dst.ctypes.data, W, H,
                                              1 if shouldAccumulate else 0, 1 if viewAccumulation else 0) == 0
