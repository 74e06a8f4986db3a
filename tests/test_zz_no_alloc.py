"""rs_render_frame_noalloc through the C ABI: frames bit-identical to rs_render_frame's on two
calls in a row, the second with a smaller frame, with the scratch taken from the library's
fixed device array; a frame whose buffers would have to grow is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

import rs_amd
from rs_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _frame(p, cam, W, H, color, noalloc):
    io = abi.FrameIO()
    io.width, io.height = W, H
    io.d_color = color.data_ptr()
    fn = p._lib.rs_render_frame_noalloc if noalloc else p._lib.rs_render_frame
    return fn(p._ctx, C.byref(cam), C.byref(io), None)


@pytest.mark.parametrize("props", [
    {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped"},
    {"temporalReuse": "GatherOnly", "gatherOption": "Robust", "neighborCount": 2, "numSpatialIterations": 2},
    {"temporalReuse": "MultiScatter", "numTimePartitions": 3},
])
def test_noalloc_matches_render_frame(props):
    lib = rs_amd.load_library()
    assert hasattr(lib, "rs_render_frame_noalloc")
    scene = scenes.box_room(seed=0)
    shutter = 1.0 / 30.0 if props["temporalReuse"] == "MultiScatter" else 0.0
    a = rs_amd.ReservoirSplatting({**props, "maxSurfaceBounces": 4})
    b = rs_amd.ReservoirSplatting({**props, "maxSurfaceBounces": 4})
    a.setScene(scene)
    b.setScene(scene)
    sizes = [(96, 64), (96, 64), (96, 64), (64, 40), (64, 40)]
    ca = torch.empty((64, 96, 4), device="cuda")
    cb = torch.empty((64, 96, 4), device="cuda")
    for f, (W, H) in enumerate(sizes):
        print(f"frame {f} {W}x{H} {props}", flush=True)
        cam = scenes.box_room_camera(f, W, H, shutter=shutter, speed=0.03)
        va, vb = ca[:H].view(-1)[:W * H * 4], cb[:H].view(-1)[:W * H * 4]
        assert _frame(a, cam, W, H, va, False) == abi.RS_OK
        # the first frame allocates the context's buffers; every later one comes through the
        # entry point that allocates nothing
        code = _frame(b, cam, W, H, vb, noalloc=f > 0)
        assert code == abi.RS_OK, lib.rs_last_error(b._ctx)
        torch.cuda.synchronize()
        assert torch.equal(va, vb), f
        for name in ("history_reservoirs", "history_recon", "color"):
            assert np.array_equal(a.read_buffer(name, W, H), b.read_buffer(name, W, H)), (f, name)
        assert a.pixelStats == b.pixelStats


def test_noalloc_refuses_growth_and_oversize():
    scene = scenes.box_room(seed=0)
    p = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterBackup", "maxSurfaceBounces": 4})
    p.setScene(scene)
    lib = p._lib
    color = torch.empty((128, 128, 4), device="cuda")
    cam = scenes.box_room_camera(0, 64, 64)
    # nothing rendered yet: the context has no buffers
    assert _frame(p, cam, 64, 64, color.view(-1)[:64 * 64 * 4], True) == abi.RS_E_SIZE_MISMATCH
    assert _frame(p, cam, 64, 64, color.view(-1)[:64 * 64 * 4], False) == abi.RS_OK
    # a larger frame would have to grow the history
    cam = scenes.box_room_camera(1, 128, 128)
    assert _frame(p, cam, 128, 128, color.view(-1), True) == abi.RS_E_SIZE_MISMATCH
    assert b"rs_render_frame" in lib.rs_last_error(p._ctx)
    # a frame whose scratch exceeds the fixed array (1024 x 1024 ScatterBackup needs ~390 MiB)
    big = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterBackup", "maxSurfaceBounces": 4})
    big.setScene(scene)
    out = torch.empty((1024, 1024, 4), device="cuda")
    cam = scenes.box_room_camera(0, 1024, 1024)
    assert _frame(big, cam, 1024, 1024, out, False) == abi.RS_OK
    assert _frame(big, cam, 1024, 1024, out, True) == abi.RS_E_SIZE_MISMATCH
    assert b"MiB" in lib.rs_last_error(big._ctx)
