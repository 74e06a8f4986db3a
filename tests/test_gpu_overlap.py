"""Frames whose stages overlap (the primary hits, Robust shifts, collect, reprojection and splat
sort on the pass's side stream while the initial candidates trace) against the same frames
with stage timing on, which runs every stage in order on the caller's stream: every buffer
and the ray counts must be identical, also on a caller stream, with the caller's VBuffer and
motion vectors, and with the per-pixel ray counts."""
import numpy as np
import pytest
import torch

import rs_amd
from rs_amd import scenes

from test_boundary_cpu import pack_vbuffer

pytestmark = pytest.mark.gpu

W, H = 160, 96
FRAME_BUFS = ["mvec", "vbuffer", "initial_reservoirs", "initial_recon", "history_reservoirs", "history_recon", "color"]
TEMPORAL_BUFS = ["temporal_reservoirs", "temporal_recon"]
COLLECT_BUFS = ["backup_reservoirs", "backup_recon", "floating_coords"]


def _written(mode, f):
    """The buffers frame f writes: temporal reuse starts at frame 1, and only GatherOnly and
    ScatterBackup collect."""
    if f == 0:
        return FRAME_BUFS
    return FRAME_BUFS + TEMPORAL_BUFS + (COLLECT_BUFS if mode in ("GatherOnly", "ScatterBackup") else [])


def _pass(props, scene, timing):
    p = rs_amd.ReservoirSplatting(dict(props, captureStages=1))
    p.setScene(scene)
    p.set_timing(timing)
    return p


@pytest.mark.parametrize("props,shutter", [
    ({"temporalReuse": "ScatterBackup", "gatherOption": "Robust"}, 0.0),
    ({"temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "neighborCount": 2}, 0.0),
    ({"temporalReuse": "GatherOnly", "gatherOption": "Fast"}, 0.0),
    ({"temporalReuse": "MultiScatter", "numTimePartitions": 4}, 1.0 / 24.0),
])
def test_overlapped_stages_equal_ordered(props, shutter):
    scene = scenes.box_room(seed=3)
    over, ordered = _pass(props, scene, False), _pass(props, scene, True)
    stream = torch.cuda.Stream()
    rc = [torch.zeros((H, W), dtype=torch.int32, device="cuda") for _ in range(2)]
    for f in range(5):
        print(f"frame {f} {props}", flush=True)
        cam = scenes.box_room_camera(f, W, H, shutter=shutter)
        with torch.cuda.stream(stream):
            over.execute(cam, W, H, stream=stream, rayCount=rc[0])
        ordered.execute(cam, W, H, rayCount=rc[1])
        stream.synchronize()
        torch.cuda.synchronize()
        for name in _written(props["temporalReuse"], f):
            assert np.array_equal(over.read_buffer(name, W, H), ordered.read_buffer(name, W, H)), (f, name)
        if f > 0 and props["temporalReuse"] != "GatherOnly":
            K = props.get("numTimePartitions", 1)
            for name in ("splat_dst", "cell_count", "cell_offset"):
                assert np.array_equal(over.read_buffer(name, W, H, K), ordered.read_buffer(name, W, H, K)), (f, name)
        assert over.pixelStats == ordered.pixelStats
        assert torch.equal(rc[0], rc[1])


def test_overlapped_stages_with_caller_vbuffer_and_mvec():
    """The caller's VBuffer is decoded and its motion vectors copied on the side stream."""
    scene = scenes.box_room(seed=4)
    props = {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped"}
    src = _pass(props, scene, True)
    over, ordered = _pass(props, scene, False), _pass(props, scene, True)
    for f in range(4):
        print(f"frame {f} caller vbuffer and mvec", flush=True)
        cam = scenes.box_room_camera(f, W, H)
        src.execute(cam, W, H)
        dvb = torch.from_numpy(pack_vbuffer(src.read_buffer("vbuffer", W, H), scene, 0).view(np.int32)).cuda()
        mv = torch.from_numpy(src.read_buffer("mvec", W, H).view(np.float32).reshape(H, W, 2).copy())
        mv = (mv * 0.5).cuda()  # not the computed vectors, so that the copy shows
        for p in (over, ordered):
            p.execute(cam, W, H, mvec=mv, vbuffer=dvb)
        torch.cuda.synchronize()
        for name in _written("ScatterBackup", f):
            assert np.array_equal(over.read_buffer(name, W, H), ordered.read_buffer(name, W, H)), (f, name)
        assert np.array_equal(over.read_buffer("mvec", W, H), mv.cpu().numpy().reshape(-1).view(np.uint8))


def test_overlapped_stages_long_run():
    """24 frames of ScatterBackup (Robust gather, two neighbours over two spatial iterations)
    on the larger frame, overlapped and ordered, compared every frame."""
    scene = scenes.box_room(seed=5)
    props = {"temporalReuse": "ScatterBackup", "gatherOption": "Robust", "neighborCount": 2,
             "numSpatialIterations": 2}
    w, h = 256, 192
    over, ordered = _pass(props, scene, False), _pass(props, scene, True)
    for f in range(24):
        cam = scenes.box_room_camera(f, w, h, aperture=0.04, focal_distance=3.5)
        over.execute(cam, w, h)
        ordered.execute(cam, w, h)
        torch.cuda.synchronize()
        for name in ("history_reservoirs", "history_recon", "color"):
            assert np.array_equal(over.read_buffer(name, w, h), ordered.read_buffer(name, w, h)), (f, name)
    print("24 frames equal", flush=True)


def test_overlapped_1080p_field_matches_ordered():
    """The benchmark workload (1080p ScatterBackup on the ~1M-triangle field): 30 frames in two
    overlapped contexts and one ordered context end in identical histories and images, so the
    overlapped stages neither race with each other nor depend on timing at full size."""
    props = {"samplesPerPixel": 1, "maxSurfaceBounces": 6, "temporalReuse": "ScatterBackup",
             "gatherOption": "Clamped", "scatterBackupMISOption": "Balance", "enableSpatialResampling": True,
             "numSpatialIterations": 1, "neighborCount": 1}
    W, H = 1920, 1080
    scene = scenes.instanced_field(seed=2)
    passes = [_pass(props, scene, timing) for timing in (False, False, True)]
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    for p in passes:
        for f in range(30):
            p.execute(scenes.orbit_camera(f, W, H), W, H, color=color)
        torch.cuda.synchronize()
        print("30 frames rendered", flush=True)
    for name in ("history_reservoirs", "history_recon", "temporal_reservoirs", "color"):
        ref = passes[2].read_buffer(name, W, H)
        for p in passes[:2]:
            assert np.array_equal(p.read_buffer(name, W, H), ref), name
