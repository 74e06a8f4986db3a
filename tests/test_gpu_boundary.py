"""The drop-in boundary's additions on the GPU, bit-exact against the oracle: rigid instance
motion (rs_update_instances), the frozen scene (rs_set_frozen), a caller VBuffer, per-pixel
ray counts and path lengths, the state checkpoint (rs_save_state / rs_load_state), the sized
replay entry points, and a failed scene build leaving no scene behind."""
import numpy as np
import pytest
import torch

import oracle_harness as oh
import rs_amd
from rs_amd import abi, scenes

from test_boundary_cpu import _rot_y, pack_vbuffer
from test_gpu_parity import _compare, _pair

pytestmark = pytest.mark.gpu


def _moving_transforms(scene, f):
    """Both boxes rotate about their centres and drift; the light slides."""
    X = np.tile(np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32), (scene.num_instances, 1))
    X[5] = _rot_y(4.0 * f, (-0.35, -0.4, -0.3), t=(0.01 * f, 0.0, 0.0))
    X[6] = _rot_y(-7.0 * f, (0.35, -0.7, 0.3), t=(0.0, 0.0, 0.015 * f))
    X[7] = _rot_y(0.0, (0, 0, 0), t=(0.02 * f, 0.0, 0.0))
    return X


@pytest.mark.parametrize("mode,shutter", [("ScatterBackup", 0.0), ("MultiScatter", 1.0 / 24.0)])
def test_instances_bit_exact(mode, shutter):
    W, H = 72, 48
    scene = scenes.box_room(seed=0)
    gpu, cpu = _pair({"temporalReuse": mode, "gatherOption": "Clamped", "numTimePartitions": 2,
                      "maxSurfaceBounces": 4}, scene)
    for f in range(5):
        print(f"frame {f} instances {mode}", flush=True)
        X = _moving_transforms(scene, f)
        gpu.updateInstances(X)
        cpu.updateInstances(X)
        cam = scenes.box_room_camera(f, W, H, shutter=shutter)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0:
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_frozen_bit_exact():
    """Frames before, during and after a frozen stretch (spatial reuse with 2 iterations, so
    the temporary ping-pong buffers carry the spatial passes), and reset() ignored while frozen."""
    W, H = 64, 48
    scene = scenes.box_room(seed=0)
    gpu, cpu = _pair({"temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "maxSurfaceBounces": 4,
                      "numSpatialIterations": 2, "neighborCount": 2}, scene)
    for f, frozen in enumerate([False, False, True, True, False, False]):
        print(f"frame {f} frozen {frozen}", flush=True)
        gpu.frozen = frozen
        cpu.frozen = frozen
        if f == 3:
            gpu.reset()
            cpu.reset()
        cam = scenes.box_room_camera(min(f, 2), W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "final_reservoirs", "final_recon",
                                  "temporal_reservoirs", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats
        assert gpu.stats().frameCount == cpu.stats().frameCount
        assert gpu.frameFlags() == cpu.frameFlags()
    assert gpu.stats().frameCount == 4


@pytest.mark.parametrize("prim_bits", [0, 14])
def test_caller_vbuffer_bit_exact(prim_bits):
    W, H = 64, 40
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": "GatherOnly", "gatherOption": "Fast", "maxSurfaceBounces": 4}
    gpu, cpu = _pair(props, scene)
    ref = oh.OraclePass(gpu._opts)
    ref.setScene(scene)
    for f in range(3):
        print(f"frame {f} caller vbuffer bits {prim_bits}", flush=True)
        cam = scenes.box_room_camera(f, W, H, aperture=0.05, focal_distance=3.5)
        ref.execute(cam, W, H)
        packed = pack_vbuffer(ref.read_buffer("vbuffer", W, H), scene, prim_bits)
        if f == 2:
            packed[5, 0] = 2 << 29  # a non-triangle hit type decodes as a miss
        dvb = torch.from_numpy(packed.view(np.int32)).cuda()
        gpu.execute(cam, W, H, vbuffer=dvb, vbufferPrimitiveIndexBits=prim_bits)
        cpu.execute(cam, W, H, vbuffer=packed, vbufferPrimitiveIndexBits=prim_bits)
        _compare(gpu, cpu, W, H, ["vbuffer", "mvec", "history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_pixel_stats_bit_exact():
    W, H = 64, 48
    scene = scenes.box_room(seed=0)
    gpu, cpu = _pair({"temporalReuse": "ScatterBackup", "maxSurfaceBounces": 4, "samplesPerPixel": 2}, scene)
    rc = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    pl = torch.full((H, W), 7, dtype=torch.int32, device="cuda")
    for f in range(3):
        print(f"frame {f} pixel stats", flush=True)
        cam = scenes.box_room_camera(f, W, H)
        gpu.execute(cam, W, H, rayCount=rc, pathLength=pl)
        cpu.execute(cam, W, H, stats=True)
        torch.cuda.synchronize()
        assert np.array_equal(rc.cpu().numpy().view(np.uint32), cpu.rayCount), f
        assert np.array_equal(pl.cpu().numpy().view(np.uint32), cpu.pathLength), f
        assert int(rc.sum()) == gpu.pixelStats["totalRays"]


def test_save_load_state_round_trip():
    """A context that loads a checkpoint renders the next frames bit for bit like the
    context that saved it."""
    W, H = 64, 48
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": "MultiScatter", "numTimePartitions": 2, "maxSurfaceBounces": 4}
    a = rs_amd.ReservoirSplatting(props)
    a.setScene(scene)
    cams = [scenes.box_room_camera(f, W, H, shutter=1.0 / 30.0, speed=0.03) for f in range(5)]
    for f in range(3):
        a.execute(cams[f], W, H)
    blob = a.saveState()
    b = rs_amd.ReservoirSplatting(props)
    b.setScene(scene)
    b.loadState(blob)
    for f in (3, 4):
        print(f"frame {f} after load", flush=True)
        a.execute(cams[f], W, H)
        b.execute(cams[f], W, H)
        for name in ("history_reservoirs", "history_recon", "color"):
            assert np.array_equal(a.read_buffer(name, W, H), b.read_buffer(name, W, H)), (f, name)
        assert a.pixelStats == b.pixelStats
    other = rs_amd.ReservoirSplatting(props)
    other.setScene(scenes.patch_room())
    with pytest.raises(rs_amd.RSError):
        other.loadState(blob)


def test_replay_sizes_and_snapshot_lifetime():
    W, H = 32, 24
    scene = scenes.box_room(seed=0)
    rs = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterOnly", "maxSurfaceBounces": 4})
    rs.setScene(scene)
    rs.execute(scenes.box_room_camera(0, W, H), W, H)
    rs.freeze()
    theta = torch.from_numpy(rs.materialParams()).cuda()
    img = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    rs.replay_forward(theta, img)
    with pytest.raises(rs_amd.RSError) as e:
        rs.replay_forward(theta[:-1].contiguous(), img)
    assert e.value.code == abi.RS_E_SIZE_MISMATCH
    with pytest.raises(rs_amd.RSError) as e:
        rs.replay_forward(theta, torch.empty((H, W - 1, 4), dtype=torch.float32, device="cuda"))
    assert e.value.code == abi.RS_E_SIZE_MISMATCH
    with pytest.raises(rs_amd.RSError):
        rs.replay_forward(theta.double(), img)
    rs.updateInstances(_moving_transforms(scene, 1))
    with pytest.raises(rs_amd.RSError) as e:
        rs.replay_forward(theta, img)
    assert e.value.code == abi.RS_E_NO_SCENE


def _chain_scene(n=95, ratio=2.5, lit=False):
    """Triangle k spans [x_k, 2 x_k] x [0, x_k] with x_k = ratio^(k - n/2): the cluster of
    triangles 0..k-1 is always the nearest neighbour of triangle k (union area about 4 x_k^2
    against 20 x_k^2 for triangle k + 1), so the clustering builds a chain n levels deep,
    about n/3 levels of 4-wide nodes: deeper than the traversal stack supports.  `lit` adds
    a floor and an emissive quad above the chain's middle, for a picture."""
    s = scenes.Scene()
    m = s.add_material((0.5, 0.5, 0.5), roughness=1.0, lobes=abi.LOBE_DIFFUSE_REFLECTION)
    for k in range(n):
        x = ratio ** (k - n // 2)
        s.add_mesh([(x, 0, 0), (2 * x, 0, 0), (x, x, 0)], [(0, 0, 1)] * 3, (0, 1, 2), m)
    if lit:
        f = s.add_material((0.6, 0.5, 0.4), roughness=1.0, lobes=abi.LOBE_DIFFUSE_REFLECTION)
        s.add_mesh([(-4, -1, -4), (8, -1, -4), (8, -1, 4), (-4, -1, 4)], [(0, 1, 0)] * 4, (0, 2, 1, 0, 3, 2), f)
        e = s.add_material((0.0, 0.0, 0.0), roughness=1.0, emissive=(15.0, 15.0, 15.0),
                           lobes=abi.LOBE_DIFFUSE_REFLECTION)
        s.add_mesh([(0, 4, -1), (2, 4, -1), (2, 4, 1), (0, 4, 1)], [(0, -1, 0)] * 4, (0, 1, 2, 0, 2, 3), e)
    return s.finalize()


def test_too_deep_clustering_tree_falls_back_to_balanced_tree():
    """A scene whose clustering tree is deeper than the traversal stack holds is built as a
    balanced tree over the Morton order instead; the picture does not depend on the tree, so
    frames stay bit-exact against the oracle."""
    W, H = 64, 48
    scene = _chain_scene(lit=True)
    props = {"temporalReuse": "ScatterBackup", "maxSurfaceBounces": 4}
    gpu = rs_amd.ReservoirSplatting(props)
    gpu.setScene(scene)
    st = gpu.stats()
    assert st.bvhBuild == abi.BVH_BUILD_BALANCED and 3 * st.bvhLevels <= 84, (st.bvhBuild, st.bvhLevels)
    gpu.setScene(scenes.box_room(seed=0))
    assert gpu.stats().bvhBuild == abi.BVH_BUILD_CLUSTERING
    gpu.setScene(scene)
    cpu = oh.OraclePass(gpu._opts)
    cpu.setScene(scene)
    for f in range(3):
        cam = scenes.make_camera((1.0 + 0.02 * f, 1.5, 5.0), (1.0, 0.5, 0.0), 45.0, W / H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        for name in ("history_reservoirs", "history_recon", "color"):
            assert np.array_equal(gpu.read_buffer(name, W, H), cpu.read_buffer(name, W, H)), (f, name)
    assert float(np.abs(gpu.read_buffer("color", W, H).view(np.float32)).sum()) > 0


def test_failed_scene_build_leaves_no_scene(monkeypatch):
    """A build that fails after the old scene is released (forced here by capping the tree
    depth below what either build reaches) leaves the context without a scene."""
    W, H = 16, 16
    rs = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterOnly", "maxSurfaceBounces": 4})
    rs.setScene(scenes.box_room(seed=0))
    color = torch.full((H, W, 4), 3.0, device="cuda")
    rs.execute(scenes.box_room_camera(0, W, H), W, H, color=color)
    monkeypatch.setenv("RS_BVH_MAX_LEVELS", "1")
    with pytest.raises(rs_amd.RSError) as e:
        rs.setScene(_chain_scene())
    assert e.value.code == abi.RS_E_UNSUPPORTED, e.value
    monkeypatch.delenv("RS_BVH_MAX_LEVELS")
    with pytest.raises(rs_amd.RSError) as e:
        rs.execute(scenes.box_room_camera(1, W, H), W, H, color=color)
    assert e.value.code == abi.RS_E_NO_SCENE
    torch.cuda.synchronize()
    assert float(color.abs().sum()) == 0.0
    rs.setScene(scenes.box_room(seed=0))
    rs.execute(scenes.box_room_camera(0, W, H), W, H, color=color)
