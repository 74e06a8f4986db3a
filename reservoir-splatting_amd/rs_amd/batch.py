"""Batched frames over the C ABI (rs_batch_*): many ReservoirSplatting passes rendered per
call as one hipGraph launch, either fused (queues 0, the default: one launch per stage for
all passes) or one side stream per queue (queues 1..8), with the pixel-tile width set by
hand or by the autotuner."""
import ctypes as C

from . import abi
from .pass_ import RSError, _dptr, load_library


class Batch:
    def __init__(self, passes):
        self._lib = load_library()
        self.passes = list(passes)
        arr = (C.c_void_p * len(self.passes))(*[p._ctx.value for p in self.passes])
        self._b = C.c_void_p()
        code = self._lib.rs_batch_create(arr, len(self.passes), C.byref(self._b))
        if code != abi.RS_OK:
            raise RSError(code, "rs_batch_create failed")

    def close(self):
        if self._b:
            self._lib.rs_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code):
        if code != abi.RS_OK:
            raise RSError(code, self._lib.rs_batch_last_error(self._b).decode())

    def configure(self, queues, tile_width):
        """The stream form with `queues` side streams (1..8), or the fused form (0)."""
        self._check(self._lib.rs_batch_configure(self._b, queues, tile_width))

    def configure_fused(self, groups, tile_width):
        """The fused form with the passes split into `groups` groups (1..8)."""
        self._check(self._lib.rs_batch_configure_fused(self._b, groups, tile_width))

    def config(self):
        """{"form": "fused" | "streams", "queues", "groups", "tile_width"}."""
        q, g, t = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.rs_batch_get_config(self._b, C.byref(q), C.byref(g), C.byref(t)))
        return {"form": "fused" if q.value == 0 else "streams", "queues": q.value, "groups": g.value,
                "tile_width": t.value}

    @staticmethod
    def _frame_args(cams, sizes, colors, ios=None):
        """ios: optional per-pass dicts of the other rs_frame_io entries (mvec, vbuffer,
        vbufferPrimitiveIndexBits, rayCount, pathLength, sceneFlags), as for
        ReservoirSplatting.execute."""
        n = len(cams)
        cam_arr = (abi.Camera * n)(*cams)
        io_arr = (abi.FrameIO * n)()
        for i, (w, h) in enumerate(sizes):
            P = int(w) * int(h)
            extra = (ios[i] if ios is not None else None) or {}
            io_arr[i].width, io_arr[i].height = int(w), int(h)
            io_arr[i].d_mvec = _dptr(extra.get("mvec"), 2 * P, ("float32",), "mvec")
            io_arr[i].d_color = _dptr(colors[i], 4 * P, ("float32",), "color") if colors is not None else None
            io_arr[i].d_vbuffer = _dptr(extra.get("vbuffer"), 4 * P, ("int32", "uint32"), "vbuffer")
            io_arr[i].vbufferPrimitiveIndexBits = int(extra.get("vbufferPrimitiveIndexBits", 0))
            io_arr[i].d_rayCount = _dptr(extra.get("rayCount"), P, ("int32", "uint32"), "rayCount")
            io_arr[i].d_pathLength = _dptr(extra.get("pathLength"), P, ("int32", "uint32"), "pathLength")
            io_arr[i].sceneFlags = int(extra.get("sceneFlags", 0))
        return cam_arr, io_arr

    def render(self, cams, sizes, colors=None, stream=None, ios=None):
        cam_arr, io_arr = self._frame_args(cams, sizes, colors, ios)
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        self._check(self._lib.rs_batch_render(self._b, cam_arr, io_arr, st))

    def autotune(self, cams, sizes, colors=None, frames=2, stream=None):
        """Returns (queues, tile width, ms per batch frame) of the fastest candidate (queues 0:
        the fused form, whose group count config() reports)."""
        cam_arr, io_arr = self._frame_args(cams, sizes, colors)
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        q, t, ms = C.c_uint32(), C.c_uint32(), C.c_float()
        self._check(self._lib.rs_batch_autotune(self._b, cam_arr, io_arr, frames, st, C.byref(q), C.byref(t),
                                                C.byref(ms)))
        return q.value, t.value, ms.value
