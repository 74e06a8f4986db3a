"""The boundary additions on the oracle (CPU): rigid instance motion, the frozen scene of
"Replay for Accumulation", a caller-supplied VBuffer and the per-pixel PixelStats outputs.
Each is pinned by a property the reference's behaviour fixes, not by retyping its formula."""
import numpy as np
import pytest

import oracle_harness as oh
from rs_amd import abi, scenes

from test_gpu_parity import BRANCHES


def _opts(props):
    o = abi.Options()
    base = dict(samplesPerPixel=1, maxSurfaceBounces=4, maxDiffuseBounces=4, maxSpecularBounces=4,
                maxTransmissionBounces=4, useRussianRoulette=1, misHeuristic=0, misPowerExponent=2.0,
                specularRoughnessThreshold=0.25, adjustShadingNormals=0, gbufferAdjustShadingNormals=1,
                enableTemporalResampling=1, temporalReuse=1, gatherMechanism=1, scatterBackupMIS=0,
                numTimePartitions=2, useConfidenceWeightsTemporally=1, enableSpatialResampling=1,
                numSpatialIterations=1, neighborCount=1, useConfidenceWeightsSpatially=1, spatialGatherRadius=30.0,
                artificialFrameTime=1.0 / 60.0, useFixedSeed=0, fixedSeed=1, captureStages=0,
                quirkMask=abi.QUIRKS_REFERENCE)
    base.update(props)
    for k, v in base.items():
        setattr(o, k, v)
    return o


def _pass(props, scene=None):
    p = oh.OraclePass(_opts(props))
    p.setScene(scene if scene is not None else scenes.box_room(seed=0))
    return p


def _rot_y(deg, center, t=(0.0, 0.0, 0.0)):
    a = np.radians(deg)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    c = np.asarray(center, np.float64)
    M = np.zeros((3, 4))
    M[:, :3] = R
    M[:, 3] = c - R @ c + np.asarray(t)
    return M.astype(np.float32).reshape(12)


def test_instance_translation_equals_moved_vertices():
    """A translation-only transform adds the offset to every coordinate exactly, so the
    instance path reproduces moved_vertices bit for bit."""
    scene = scenes.box_room(seed=0)
    vel = {5: (0.03, 0.0, 0.01), 7: (0.01, 0.0, 0.0)}
    p = _pass({}, scene)
    p.updateInstances(scenes.instance_transforms(scene, 3, vel))
    P, _ = p.vertices()
    assert np.array_equal(P, scenes.moved_vertices(scene, 3, vel))


def test_instance_rotation_is_rigid_and_normals_follow():
    """A rotation about the box's centre keeps every distance within the instance and maps
    normals by the inverse transpose (here the rotation itself): compared with float64."""
    scene = scenes.box_room(seed=0)
    X = np.tile(np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32), (scene.num_instances, 1))
    X[6] = _rot_y(37.0, (0.35, -0.7, 0.3), t=(0.05, 0.02, -0.04))
    # a mirroring, scaling transform on a wall: normals by the inverse transpose, flipped with det < 0
    S = np.diag([-1.5, 0.5, 2.0])
    X[2, [0, 1, 2, 4, 5, 6, 8, 9, 10]] = S.reshape(-1)
    p = _pass({}, scene)
    p.updateInstances(X)
    P, N = p.vertices()
    fv, nv = scene.mesh_vertices[6]
    M = X[6].reshape(3, 4).astype(np.float64)
    obj = scene.positions[fv:fv + nv].astype(np.float64)
    ref = obj @ M[:, :3].T + M[:, 3]
    np.testing.assert_allclose(P[fv:fv + nv], ref, rtol=0, atol=2e-6)
    d_obj = np.linalg.norm(obj[:, None] - obj[None], axis=-1)
    d_new = np.linalg.norm(P[fv:fv + nv, None].astype(np.float64) - P[None, fv:fv + nv], axis=-1)
    np.testing.assert_allclose(d_new, d_obj, atol=4e-6)
    nref = scene.normals[fv:fv + nv].astype(np.float64) @ M[:, :3].T
    np.testing.assert_allclose(N[fv:fv + nv], nref / np.linalg.norm(nref, axis=1, keepdims=True), atol=2e-6)
    fv2, nv2 = scene.mesh_vertices[2]
    A = X[2].reshape(3, 4)[:, :3].astype(np.float64)
    nref = scene.normals[fv2:fv2 + nv2].astype(np.float64) @ np.linalg.inv(A)  # (A^-T n) as rows
    nref /= np.linalg.norm(nref, axis=1, keepdims=True)
    np.testing.assert_allclose(N[fv2:fv2 + nv2], nref, atol=2e-6)
    # untouched instances keep their vertices
    fv0, nv0 = scene.mesh_vertices[0]
    assert np.array_equal(P[fv0:fv0 + nv0], scene.positions[fv0:fv0 + nv0])


def test_frozen_frames_keep_history_and_rejoin():
    """While frozen, a frame neither advances frameCount nor touches the history, reset() is
    ignored, and the dictionary suppresses auto-reset; with a fixed seed a frozen re-render is
    identical to the frame it repeats, and after unfreezing the sequence continues exactly as
    in a context that never froze (RS/ReservoirSplatting.cpp:503-506, 813-820, 1672-1679,
    2003-2004)."""
    W, H = 40, 32
    props = {"temporalReuse": 2, "useFixedSeed": 1, "fixedSeed": 3, "numSpatialIterations": 2}
    a, b = _pass(props), _pass(props)
    for f in range(3):
        cam = scenes.box_room_camera(f, W, H)
        a.execute(cam, W, H)
        b.execute(cam, W, H)
    hist = b.read_buffer("history_reservoirs", W, H).copy()
    histC = b.read_buffer("history_recon", W, H).copy()
    b.frozen = True
    cam3 = scenes.box_room_camera(3, W, H)
    b.execute(cam3, W, H)
    first = b.color.copy()
    assert b.stats().frameCount == 3
    assert b.frameFlags() == {"_suppressAutoReset": True, "_shouldAccumulate": True, "_viewAccumulation": False}
    b.reset()
    b.execute(cam3, W, H)
    assert np.array_equal(b.color, first)
    assert np.array_equal(b.read_buffer("history_reservoirs", W, H), hist)
    assert np.array_equal(b.read_buffer("history_recon", W, H), histC)
    assert b.stats().frameCount == 3
    b.frozen = False
    for f in (3, 4):
        cam = scenes.box_room_camera(f, W, H)
        a.execute(cam, W, H)
        b.execute(cam, W, H)
        assert np.array_equal(a.color, b.color), f
        assert np.array_equal(a.read_buffer("history_recon", W, H), b.read_buffer("history_recon", W, H)), f
    assert b.frameFlags()["_suppressAutoReset"] is False


def test_frozen_replay_equals_unfrozen_frame():
    """With a fixed seed, the frozen render of frame k equals the ordinary frame k."""
    W, H = 40, 32
    props = {"temporalReuse": 1, "useFixedSeed": 1, "fixedSeed": 5}
    a, b = _pass(props), _pass(props)
    for f in range(3):
        cam = scenes.box_room_camera(f, W, H)
        a.execute(cam, W, H)
        if f < 2:
            b.execute(cam, W, H)
    b.frozen = True
    b.execute(scenes.box_room_camera(2, W, H), W, H)
    assert np.array_equal(a.color, b.color)
    assert np.array_equal(a.read_buffer("history_recon", W, H), b.read_buffer("final_recon", W, H))


def test_scene_flags_dictionary():
    W, H = 16, 12
    p = _pass({})
    cam = scenes.box_room_camera(0, W, H)
    p.execute(cam, W, H, sceneFlags=abi.SCENE_REPLAYING)
    assert p.frameFlags() == {"_suppressAutoReset": True, "_shouldAccumulate": False, "_viewAccumulation": False}
    p.execute(cam, W, H, sceneFlags=abi.SCENE_REPLAYING | abi.SCENE_ACCUMULATION_FRAME | abi.SCENE_VIEW_ACCUMULATION)
    assert p.frameFlags() == {"_suppressAutoReset": True, "_shouldAccumulate": True, "_viewAccumulation": True}
    p.execute(cam, W, H)
    assert p.frameFlags() == {"_suppressAutoReset": False, "_shouldAccumulate": True, "_viewAccumulation": False}


def test_scene_should_reset_skips_temporal_only():
    """Scene::shouldReset skips temporal reuse and refills the keyframes but keeps frameCount."""
    W, H = 32, 24
    p = _pass({"temporalReuse": 1})
    for f in range(3):
        p.execute(scenes.box_room_camera(f, W, H), W, H)
    before = p.pixelStats["totalRays"]
    p.execute(scenes.box_room_camera(3, W, H), W, H, sceneFlags=abi.SCENE_SHOULD_RESET)
    assert p.stats().frameCount == 4
    q = _pass({"temporalReuse": 1, "enableTemporalResampling": 0})
    for f in range(4):
        q.execute(scenes.box_room_camera(f, W, H), W, H)
    assert before > 0 and p.pixelStats["totalRays"] == q.pixelStats["totalRays"]


def pack_vbuffer(vb, scene, prim_bits=0):
    """PackedHitInfo (4 words per pixel) of a (prim, u, v) VBuffer, instance = mesh."""
    w = np.frombuffer(vb.tobytes(), np.uint32).reshape(-1, 3)
    prim = w[:, 0]
    out = np.zeros((w.shape[0], 4), np.uint32)
    hit = prim != 0xFFFFFFFF
    starts = np.array([t for t, _ in scene.mesh_triangles], np.int64)
    inst = np.searchsorted(starts, prim[hit].astype(np.int64), side="right") - 1
    local = prim[hit].astype(np.int64) - starts[inst]
    if prim_bits == 0:
        out[hit, 0] = (1 << 29) | inst
        out[hit, 1] = local
    else:
        out[hit, 0] = (1 << 29) | (inst << prim_bits) | local
    out[:, 2:] = w[:, 1:]
    return out


@pytest.mark.parametrize("prim_bits", [0, 12])
def test_caller_vbuffer_reproduces_traced_frame(prim_bits):
    """A VBuffer packed from the pass's own primary hits renders the identical frame; garbage
    entries (bad type, instance or primitive) decode as misses."""
    W, H = 48, 32
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": 0, "gatherMechanism": 0}
    a, b = _pass(props, scene), _pass(props, scene)
    for f in range(3):
        cam = scenes.box_room_camera(f, W, H, aperture=0.05, focal_distance=3.5)
        a.execute(cam, W, H)
        packed = pack_vbuffer(a.read_buffer("vbuffer", W, H), scene, prim_bits)
        b.execute(cam, W, H, vbuffer=packed, vbufferPrimitiveIndexBits=prim_bits)
        for name in ("vbuffer", "mvec", "history_reservoirs", "history_recon", "color"):
            assert np.array_equal(a.read_buffer(name, W, H), b.read_buffer(name, W, H)), (f, name)
    bad = pack_vbuffer(a.read_buffer("vbuffer", W, H), scene, prim_bits)
    bad[0, 0] = (2 << 29)                       # a volume hit
    bad[1, 0] = (1 << 29) | (999 << prim_bits if prim_bits else 999)  # no such instance
    if prim_bits == 0:
        bad[2, 1] = 10 ** 6                     # no such primitive
    b.execute(scenes.box_room_camera(3, W, H), W, H, vbuffer=bad, vbufferPrimitiveIndexBits=prim_bits)
    vb = np.frombuffer(b.read_buffer("vbuffer", W, H).tobytes(), np.uint32).reshape(-1, 3)
    assert vb[0, 0] == 0xFFFFFFFF and vb[1, 0] == 0xFFFFFFFF
    if prim_bits == 0:
        assert vb[2, 0] == 0xFFFFFFFF


def test_pixel_stats_outputs():
    """rayCount sums to the frame's ray total; without temporal and spatial reuse every pixel
    holds its own initial-candidate rays; with reuse pixel (0,0) gains all the resampling
    rays; the path length is never logged, so it is cleared."""
    W, H = 40, 32
    p = _pass({"enableTemporalResampling": 0, "enableSpatialResampling": 0})
    p.execute(scenes.box_room_camera(0, W, H), W, H, stats=True)
    own = p.rayCount.copy()
    assert int(own.sum()) == p.pixelStats["totalRays"]
    assert np.all(p.pathLength == 0)
    q = _pass({"temporalReuse": 2})
    for f in range(2):
        q.execute(scenes.box_room_camera(f, W, H), W, H, stats=True)
    assert int(q.rayCount.astype(np.int64).sum()) == q.pixelStats["totalRays"]
    r = _pass({"temporalReuse": 2, "enableTemporalResampling": 0, "enableSpatialResampling": 0})
    for f in range(2):
        r.execute(scenes.box_room_camera(f, W, H), W, H, stats=True)
    # same seeds and initial candidates: every pixel but (0,0) holds the same count
    assert np.array_equal(q.rayCount.ravel()[1:], r.rayCount.ravel()[1:])
    assert q.rayCount[0, 0] > r.rayCount[0, 0]


@pytest.mark.parametrize("name,props,aperture", BRANCHES, ids=[b[0] for b in BRANCHES])
def test_oracle_option_branches_run(name, props, aperture):
    """Every option branch the GPU parity tests cover runs on the oracle with finite output
    and consistent ray accounting (the GPU tests then compare every bit)."""
    import rs_amd
    W, H = 24, 20
    p = rs_amd.ReservoirSplatting({**props, "maxSurfaceBounces": 4})  # the product's property parsing (no GPU)
    cpu = oh.OraclePass(p._opts)
    cpu.setScene(scenes.box_room(seed=0))
    multi = props.get("temporalReuse") == "MultiScatter"
    for f in range(3):
        cam = scenes.box_room_camera(f, W, H, aperture=aperture, focal_distance=3.5,
                                     shutter=1.0 / 24.0 if multi else 0.0, speed=0.05 if multi else 0.01)
        cpu.execute(cam, W, H, stats=True)
        assert np.all(np.isfinite(cpu.color))
        assert int(cpu.rayCount.astype(np.int64).sum()) == cpu.pixelStats["totalRays"]


@pytest.mark.parametrize("make", [lambda: scenes.box_room(seed=0), lambda: scenes.patch_room(),
                                  lambda: scenes.instanced_field(seed=2, num_spheres=20, num_boxes=50),
                                  lambda: scenes.moving_field(seed=3, num_moving=4, num_static_spheres=5,
                                                              num_static_boxes=10)[0],
                                  lambda: scenes.furnace_cube()])
def test_scene_instance_tables_are_valid(make):
    """Every workload scene's instance table passes rs_set_scene's checks (disjoint vertex
    ranges, triangles indexing only their own instance), which run before any GPU call: on a
    machine without a GPU the call then fails at the device, not at the validation."""
    import ctypes as C
    import rs_amd
    scene = make()
    p = rs_amd.ReservoirSplatting({})
    code = p._lib.rs_set_scene(p._ctx, C.byref(scene.desc()))
    msg = p._lib.rs_last_error(p._ctx).decode()
    assert code != abi.RS_E_INVALID_ARG, msg
    covered = sum(n for _, n in scene.mesh_triangles)
    assert covered == scene.num_triangles
