"""AccumulatePass over the C ABI (rs_accum_*): temporal accumulation of the pass's
RGBA32F output, the step after the path.

Keeps the reference pass's property keys and behaviour
(Source/RenderPasses/AccumulatePass/AccumulatePass.cpp:57-94 properties, :130-290
execute/accumulate, :361-385 setScene/setEnabled/reset; defaults AccumulatePass.h:140-148):
`enabled`, `autoReset`, `precisionMode` ("Single", "SingleCompensated", "Double"),
`maxFrameCount` (0 = no limit) and `overflowMode` ("Stop", "Reset", "EMA"); the output
size and format keys are accepted and inert (the output is the input's size, RGBA32F).
`enableAccumulation` is the deprecated alias of `enabled`.  Unknown keys are warnings.

The reference resets automatically on scene and camera changes it learns from Falcor's
refresh flags; here the caller reports them: `execute(..., sceneChanged=True)` resets when
`autoReset` is on, and `setScene()` always resets.
"""
import ctypes as C

from . import abi
from .pass_ import RSError, load_library

_INERT = {"outputFormat", "outputSize", "fixedOutputSize"}


class AccumulatePass:
    def __init__(self, props=None):
        self._lib = load_library()
        self._opts = abi.AccumOptions()
        self._lib.rs_accum_default_options(C.byref(self._opts))
        self.autoReset = True
        self.warnings = []
        props = dict(props or {})
        if "enableAccumulation" in props:
            self.warnings.append("'enableAccumulation' is deprecated. Use 'enabled' instead.")
            if "enabled" not in props:
                props["enabled"] = props["enableAccumulation"]
            del props["enableAccumulation"]
        for key, value in props.items():
            if key == "enabled":
                self._opts.enabled = 1 if value else 0
            elif key == "autoReset":
                self.autoReset = bool(value)
            elif key == "precisionMode":
                self._opts.precision = self._enum(value, abi.ACCUM_PRECISION, key)
            elif key == "maxFrameCount":
                self._opts.maxFrameCount = int(value)
            elif key == "overflowMode":
                self._opts.overflowMode = self._enum(value, abi.ACCUM_OVERFLOW, key)
            elif key in _INERT:
                pass
            else:
                self.warnings.append(f"Unknown property '{key}' in AccumulatePass properties.")
        self._h = C.c_void_p()
        self._check(self._lib.rs_accum_create(C.byref(self._opts), C.byref(self._h)))

    @staticmethod
    def _enum(value, table, name):
        if isinstance(value, str):
            if value not in table:
                raise RSError(abi.RS_E_INVALID_ARG, f"unknown {name} '{value}'")
            return table[value]
        return int(value)

    def _check(self, code):
        if code != 0:
            msg = self._lib.rs_accum_last_error(self._h).decode() if getattr(self, "_h", None) else ""
            raise RSError(code, msg)

    def __del__(self):
        try:
            if self._h:
                self._lib.rs_accum_destroy(self._h)
        except Exception:
            pass

    def getProperties(self):
        inv = lambda t, v: next(k for k, x in t.items() if x == v)  # noqa: E731
        return {"enabled": bool(self._opts.enabled), "autoReset": self.autoReset,
                "precisionMode": inv(abi.ACCUM_PRECISION, self._opts.precision),
                "maxFrameCount": int(self._opts.maxFrameCount),
                "overflowMode": inv(abi.ACCUM_OVERFLOW, self._opts.overflowMode)}

    @property
    def enabled(self):
        return bool(self._opts.enabled)

    @enabled.setter
    def enabled(self, value):
        if bool(value) != self.enabled:  # setEnabled resets on a change
            self._opts.enabled = 1 if value else 0
            self._check(self._lib.rs_accum_set_options(self._h, C.byref(self._opts)))

    @property
    def frameCount(self):
        return int(self._lib.rs_accum_frame_count(self._h))

    def reset(self):
        self._check(self._lib.rs_accum_reset(self._h))

    def setScene(self, scene=None):
        self.reset()

    def execute(self, src, dst, stream=None, shouldAccumulate=True, viewAccumulation=False, sceneChanged=False):
        """Accumulate `src` into the running average written to `dst`: both RGBA32F device
        tensors of shape (H, W, 4)."""
        if sceneChanged and self.autoReset:
            self.reset()
        import torch
        H, W = int(src.shape[0]), int(src.shape[1])
        if tuple(src.shape) != (H, W, 4) or tuple(dst.shape) != (H, W, 4):
            raise RSError(abi.RS_E_SIZE_MISMATCH, "src and dst must be (H, W, 4) RGBA32F images of one size")
        for t in (src, dst):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise RSError(abi.RS_E_INVALID_ARG, "src and dst must be contiguous float32 device tensors")
        st = None if stream is None else C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
        self._check(self._lib.rs_accum_execute(self._h, C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), W, H,
                                               1 if shouldAccumulate else 0, 1 if viewAccumulation else 0, st))
