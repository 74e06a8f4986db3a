"""Bit-exact parity with the oracle at the workloads' full sizes.

* The 256 x 256 box room, ScatterOnly, max depth 4, seed 0, over all 16 frames of its
  translating camera.
* The 1080p ScatterBackup frame on the ~1M-triangle instanced field (Clamped gather, Balance
  MIS, max depth 6, spatial reuse), 3 frames of the orbit.
* The 1080p MultiScatter frame on ~2M triangles, 256 rigidly moving icospheres among 0.7M
  static triangles: 8 time partitions, shutter 1/24 s, simulated frame time 1/48 s, lens
  aperture 0.02 focused at 10.  The GPU refits its BVH every frame; the oracle rebuilds
  its own.
* The same 1080p field frame in the other temporal modes: GatherOnly with the Robust
  gather, GatherOnly with a lens, ScatterOnly with two spatial neighbours, and
  ScatterBackup with Pairwise MIS under a sky map.

At full size the splat lists are also checked against what the sort must produce: each
cell's count is the number of splats aimed at it, the offsets are their exclusive scan,
and each segment lists its source pixels in ascending order.  The oracle runs on all host
cores (OpenMP)."""
import numpy as np
import pytest

import oracle_harness as oh
import rs_amd
from rs_amd import scenes

pytestmark = pytest.mark.gpu

NO_SPLAT = 0xFFFFFFFF


def _pair(props, scene):
    gpu = rs_amd.ReservoirSplatting(props)
    gpu.setScene(scene)
    oh.set_threads(0)
    cpu = oh.OraclePass(gpu._opts)
    cpu.setScene(scene)
    return gpu, cpu


def _compare(gpu, cpu, W, H, names, frame):
    for name in names:
        g = gpu.read_buffer(name, W, H)
        c = cpu.read_buffer(name, W, H)
        bpp = rs_amd.abi.BUF_BYTES_PER_PIXEL[rs_amd.abi.BUF[name]]
        mm = oh.first_mismatch(g, c, W, bpp)
        assert mm is None, f"frame {frame} {name}: first mismatch (pixel, word, xy, gpu, cpu, #pixels) = {mm}"


def _check_splat_lists(gpu, cpu, W, H, K, frame):
    """The GPU's splat lists equal the oracle's and the sort's defining properties."""
    P = W * H
    dst = gpu.read_buffer("splat_dst", W, H, K).view(np.uint32)[:K * P]
    count = gpu.read_buffer("cell_count", W, H, K).view(np.uint32)[:K * P]
    offset = gpu.read_buffer("cell_offset", W, H, K).view(np.uint32)[:K * P]
    for name, g in (("splat_dst", dst), ("cell_count", count), ("cell_offset", offset)):
        c = cpu.read_buffer(name, W, H, K).view(np.uint32)[:K * P]
        assert np.array_equal(g, c), f"frame {frame} {name} differs from the oracle"
    valid = dst != NO_SPLAT
    total = int(valid.sum())
    assert total > 0
    assert np.array_equal(count, np.bincount(dst[valid], minlength=K * P).astype(np.uint32))
    assert offset[0] == 0 and np.array_equal(offset[1:], np.cumsum(count, dtype=np.uint64)[:-1].astype(np.uint32))
    src = (np.nonzero(valid)[0] % P).astype(np.uint32)
    order = np.lexsort((src, dst[valid]))
    sorted_ = gpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total]
    assert np.array_equal(sorted_, src[order]), f"frame {frame}: splat lists are not the sorted sources"
    assert np.array_equal(sorted_, cpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total])
    return total


def test_box_room_256_sixteen_frames():
    W, H = 256, 256
    scene = scenes.box_room(seed=0)
    gpu, cpu = _pair({"temporalReuse": "ScatterOnly", "maxSurfaceBounces": 4}, scene)
    for f in range(16):
        print(f"box room 256x256 frame {f}", flush=True)
        cam = scenes.box_room_camera(f, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_field_1080p_scatter_backup():
    W, H = 1920, 1080
    scene = scenes.instanced_field(seed=2)
    props = {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "maxSurfaceBounces": 6,
             "maxDiffuseBounces": 6, "maxSpecularBounces": 6, "maxTransmissionBounces": 6,
             "enableSpatialResampling": 1, "numSpatialIterations": 1, "neighborCount": 1, "captureStages": 1}
    gpu, cpu = _pair(props, scene)
    for f in range(3):
        print(f"field 1080p ScatterBackup frame {f}", flush=True)
        cam = scenes.orbit_camera(f, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0:
            n = _check_splat_lists(gpu, cpu, W, H, 1, f)
            print(f"  {n} splats", flush=True)
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_moving_field_1080p_multi_scatter():
    W, H, K = 1920, 1080, 8
    scene, vel = scenes.moving_field(seed=3)
    props = {"temporalReuse": "MultiScatter", "numTimePartitions": K, "artificialFrameTime": 1.0 / 48.0,
             "maxSurfaceBounces": 6, "maxDiffuseBounces": 6, "maxSpecularBounces": 6, "maxTransmissionBounces": 6,
             "enableSpatialResampling": 1, "numSpatialIterations": 1, "neighborCount": 1}
    gpu, cpu = _pair(props, scene)
    for f in range(3):
        print(f"moving field 1080p MultiScatter K={K} frame {f}", flush=True)
        P = scenes.moved_vertices(scene, f, vel)
        gpu.updateVertices(P, scene.normals)
        cpu.updateVertices(P, scene.normals)
        cam = scenes.orbit_camera(f, W, H, aperture=0.02, focal_distance=10.0, shutter=1.0 / 24.0)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0:
            n = _check_splat_lists(gpu, cpu, W, H, K, f)
            print(f"  {n} splats over {K} partitions", flush=True)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


_FIELD_DEPTH = {"maxSurfaceBounces": 6, "maxDiffuseBounces": 6, "maxSpecularBounces": 6, "maxTransmissionBounces": 6}


# The other temporal modes on the same full-size frame: GatherOnly with the Robust gather
# (the 8-neighbour shift ring) and with a lens (its second GRIS), ScatterOnly with two
# spatial neighbours, and ScatterBackup with Pairwise MIS lit also by the sky map.
@pytest.mark.parametrize("name,props,cam_kw,sky", [
    ("gather-robust", {"temporalReuse": "GatherOnly", "gatherOption": "Robust"}, {}, False),
    ("gather-clamped-lens", {"temporalReuse": "GatherOnly", "gatherOption": "Clamped"},
     {"aperture": 0.02, "focal_distance": 10.0}, False),
    ("scatter-only-2-neighbours", {"temporalReuse": "ScatterOnly", "neighborCount": 2}, {}, False),
    ("backup-pairwise-sky", {"temporalReuse": "ScatterBackup", "gatherOption": "Fast",
                             "scatterBackupMISOption": "Pairwise"}, {}, True),
])
def test_field_1080p_other_modes(name, props, cam_kw, sky):
    W, H = 1920, 1080
    scene = scenes.instanced_field(seed=2)
    props = dict(props, **_FIELD_DEPTH)
    gpu, cpu = _pair(props, scene)
    if sky:
        env = scenes.sky_env_map(512, 256)
        gpu.setEnvMap(env, intensity=1.0)
        cpu.setEnvMap(env, intensity=1.0)
    for f in range(3):
        print(f"field 1080p {name} frame {f}", flush=True)
        cam = scenes.orbit_camera(f, W, H, **cam_kw)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0 and props["temporalReuse"] != "GatherOnly":
            _check_splat_lists(gpu, cpu, W, H, 1, f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats
