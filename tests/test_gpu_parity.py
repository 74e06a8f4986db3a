"""GPU parity against the CPU oracle: every reservoir and reconnection bit, splat targets,
sort order and ray counts must be identical (shared seeded RNG, deterministic sort)."""
import numpy as np
import pytest

import oracle_harness as oh
import rs_amd
from rs_amd import scenes

pytestmark = pytest.mark.gpu


def _pair(props, scene):
    gpu = rs_amd.ReservoirSplatting(props)
    gpu.setScene(scene)
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


@pytest.mark.parametrize("mode,gather", [("ScatterOnly", "Clamped"), ("ScatterBackup", "Clamped"),
                                         ("ScatterBackup", "Fast"), ("GatherOnly", "Fast"), ("GatherOnly", "Clamped"),
                                         ("GatherOnly", "Robust"), ("ScatterBackup", "Robust")])
def test_box_room_bit_exact(mode, gather):
    W, H = 96, 64
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": mode, "gatherOption": gather, "maxSurfaceBounces": 4, "captureStages": 1}
    gpu, cpu = _pair(props, scene)
    for f in range(4):
        print(f"frame {f} mode {mode} gather {gather}", flush=True)
        cam = scenes.box_room_camera(f, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["mvec", "vbuffer", "initial_reservoirs", "initial_recon"], f)
        if f > 0 and mode != "GatherOnly":
            _compare(gpu, cpu, W, H, ["splat_dst", "cell_count", "cell_offset"], f)
            total = int(cpu.read_buffer("cell_count", W, H).view(np.uint32).sum())
            gs = gpu.read_buffer("sorted", W, H).view(np.uint32)[:total]
            cs = cpu.read_buffer("sorted", W, H).view(np.uint32)[:total]
            assert np.array_equal(gs, cs), f"frame {f}: sort order differs"
            assert np.array_equal(gpu.scatterCounts(W, H), cpu.scatterCounts(W, H))
        if f > 0:
            if mode != "ScatterOnly":
                _compare(gpu, cpu, W, H, ["backup_reservoirs", "backup_recon", "floating_coords"], f)
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


@pytest.mark.parametrize("K,aperture", [(2, 0.0), (4, 0.05)])
def test_multi_scatter_bit_exact(K, aperture):
    """MultiScatter under a moving camera with motion blur (and depth of field)."""
    W, H = 80, 56
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": "MultiScatter", "numTimePartitions": K, "maxSurfaceBounces": 4}
    gpu, cpu = _pair(props, scene)
    for f in range(4):
        print(f"frame {f} MultiScatter K={K} aperture={aperture}", flush=True)
        cam = scenes.box_room_camera(f, W, H, aperture=aperture, focal_distance=3.5, shutter=1.0 / 24.0, speed=0.05)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0:
            for name in ["splat_dst", "cell_count", "cell_offset"]:
                g = gpu.read_buffer(name, W, H, K)
                c = cpu.read_buffer(name, W, H, K)
                mm = oh.first_mismatch(g, c, W, 4)
                assert mm is None, f"frame {f} {name}: first mismatch {mm}"
            total = int(cpu.read_buffer("cell_count", W, H, K).view(np.uint32).sum())
            gs = gpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total]
            cs = cpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total]
            assert np.array_equal(gs, cs), f"frame {f}: sort order differs"
            assert np.array_equal(gpu.scatterCounts(W, H), cpu.scatterCounts(W, H))
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


@pytest.mark.parametrize("mode,shutter", [("ScatterBackup", 0.0), ("MultiScatter", 1.0 / 24.0)])
def test_moving_geometry_bit_exact(mode, shutter):
    """Both boxes and the light move every frame: updateVertices refits the GPU BVH and
    rebuilds the light collection (the oracle rebuilds everything); every frame stays
    bit-exact, also across a full rebuild in the middle."""
    W, H = 72, 48
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": mode, "gatherOption": "Clamped", "numTimePartitions": 2, "maxSurfaceBounces": 4}
    gpu, cpu = _pair(props, scene)
    vel = {5: (0.03, 0.0, 0.01), 6: (-0.02, 0.01, 0.0), 7: (0.01, 0.0, 0.0)}  # two boxes and the light
    for f in range(5):
        print(f"frame {f} moving {mode}", flush=True)
        P = scenes.moved_vertices(scene, f, vel)
        gpu.updateVertices(P, scene.normals)
        cpu.updateVertices(P, scene.normals)
        if f == 3:
            gpu.rebuildBvh()
        cam = scenes.box_room_camera(f, W, H, shutter=shutter)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        if f > 0:
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


@pytest.mark.parametrize("mode,gather,K,aperture", [("GatherOnly", "Robust", 2, 0.0), ("ScatterBackup", "Robust", 2, 0.0),
                                                    ("MultiScatter", "Clamped", 4, 0.05)])
def test_corrected_quirks_bit_exact(mode, gather, K, aperture):
    """quirkMask = 0 (the Robust grey integrand, the primary-hit time and MultiReproject's
    lens-less projection corrected) stays bit-exact with the oracle."""
    W, H = 80, 56
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": mode, "gatherOption": gather, "numTimePartitions": K, "maxSurfaceBounces": 4,
             "quirkMask": 0, "captureStages": 1}
    gpu, cpu = _pair(props, scene)
    for f in range(4):
        print(f"frame {f} {mode} {gather} quirks corrected", flush=True)
        cam = scenes.box_room_camera(f, W, H, aperture=aperture, focal_distance=3.5, shutter=1.0 / 24.0, speed=0.05)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["initial_reservoirs", "initial_recon"], f)
        if f > 0:
            if mode != "MultiScatter":
                _compare(gpu, cpu, W, H, ["backup_reservoirs", "backup_recon"], f)
            else:
                for name in ["splat_dst", "cell_count"]:
                    assert np.array_equal(gpu.read_buffer(name, W, H, K), cpu.read_buffer(name, W, H, K)), name
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


@pytest.mark.parametrize("mode,spp,quirks", [("ScatterBackup", 3, rs_amd.abi.QUIRKS_REFERENCE),
                                             ("GatherOnly", 2, rs_amd.abi.QUIRKS_REFERENCE),
                                             ("ScatterBackup", 4, rs_amd.abi.QUIRKS_REFERENCE
                                              & ~rs_amd.abi.QUIRK["sppLastSample"])])
def test_several_samples_per_pixel_bit_exact(mode, spp, quirks):
    """samplesPerPixel > 1: each sample's candidate is merged into the pixel's reservoir
    (addCandidateReservoir), its reconnection record staged until the merge keeps it; with the
    reference's start at the last sample, and corrected (RS_QUIRK_SPP_LAST_SAMPLE cleared)."""
    W, H = 64, 48
    scene = scenes.box_room(seed=0)
    props = {"temporalReuse": mode, "samplesPerPixel": spp, "maxSurfaceBounces": 4, "captureStages": 1,
             "quirkMask": quirks}
    gpu, cpu = _pair(props, scene)
    for f in range(3):
        print(f"frame {f} {mode} spp {spp}", flush=True)
        cam = scenes.box_room_camera(f, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["initial_reservoirs", "initial_recon", "history_reservoirs", "history_recon",
                                  "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


# Option branches beyond the defaults, each bit-exact against the oracle on every buffer the
# frame writes: the DoF second GRIS of the gather (RS/GatherTemporalResampling.rt.slang:166-211)
# and the DoF shifts of ScatterBackup under aperture, the Pairwise backup MIS
# (RS/ScatterBackupTemporalResampling.rt.slang:87, 138, 194, 255), the Power2 / PowerExp MIS
# heuristics (RS/PathTracer.slang:480-508), several spatial neighbours and iterations with the
# +1337 * iteration seed (RS/SpatialResampling.rt.slang:72, RS/ReservoirSplatting.cpp:2006-2019),
# the fixed seed (RS/InitialCandidates.cs.slang:84 and every pass's seed), confidence weights
# off and Russian roulette off.
BRANCHES = [
    ("gather-aperture-fast", {"temporalReuse": "GatherOnly", "gatherOption": "Fast"}, 0.05),
    ("gather-aperture-clamped", {"temporalReuse": "GatherOnly", "gatherOption": "Clamped"}, 0.05),
    ("gather-aperture-robust", {"temporalReuse": "GatherOnly", "gatherOption": "Robust"}, 0.05),
    ("backup-aperture-clamped", {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped"}, 0.05),
    ("backup-aperture-fast", {"temporalReuse": "ScatterBackup", "gatherOption": "Fast"}, 0.05),
    ("backup-pairwise", {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped",
                         "scatterBackupMISOption": "Pairwise"}, 0.0),
    ("backup-pairwise-aperture", {"temporalReuse": "ScatterBackup", "gatherOption": "Fast",
                                  "scatterBackupMISOption": "Pairwise"}, 0.05),
    ("mis-power2", {"temporalReuse": "ScatterBackup", "misHeuristic": "PowerTwo"}, 0.0),
    ("mis-powerexp", {"temporalReuse": "ScatterOnly", "misHeuristic": "PowerExp", "misPowerExponent": 3.0}, 0.0),
    ("spatial-3x2", {"temporalReuse": "ScatterBackup", "neighborCount": 3, "numSpatialIterations": 2}, 0.0),
    ("spatial-3x2-aperture", {"temporalReuse": "GatherOnly", "neighborCount": 3, "numSpatialIterations": 2}, 0.05),
    ("spatial-0-neighbours", {"temporalReuse": "ScatterOnly", "neighborCount": 0}, 0.0),
    ("fixed-seed", {"temporalReuse": "ScatterBackup", "fixedSeed": 7}, 0.0),
    ("confidence-off", {"temporalReuse": "ScatterBackup", "useConfidenceWeightsTemporally": 0,
                        "useConfidenceWeightsSpatially": 0}, 0.0),
    ("rr-off", {"temporalReuse": "ScatterOnly", "useRussianRoulette": 0}, 0.0),
    ("multi-3-partitions-2-neighbours", {"temporalReuse": "MultiScatter", "numTimePartitions": 3, "neighborCount": 2}, 0.05),
]


@pytest.mark.parametrize("name,props,aperture", BRANCHES, ids=[b[0] for b in BRANCHES])
def test_option_branches_bit_exact(name, props, aperture):
    W, H = 72, 56
    scene = scenes.box_room(seed=0)
    multi = props.get("temporalReuse") == "MultiScatter"
    shutter = 1.0 / 24.0 if multi else 0.0
    gpu, cpu = _pair({**props, "maxSurfaceBounces": 4, "captureStages": 1}, scene)
    for f in range(4):
        print(f"frame {f} {name}", flush=True)
        cam = scenes.box_room_camera(f, W, H, aperture=aperture, focal_distance=3.5, shutter=shutter,
                                     speed=0.05 if multi else 0.01)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["vbuffer", "mvec", "initial_reservoirs", "initial_recon"], f)
        if f > 0:
            _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon"], f)
        _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_glossy_patch_room_bit_exact():
    """The patch room (GGX specular + Frostbite diffuse, roughness U[0.2, 0.9]) exercises the
    lobe-masked reconnection with specular component types and the roughness-threshold
    bounce classification (RS/PathTracer.slang:331-382, 407-409; RS/ShiftMapping.slang:454)."""
    W, H = 72, 56
    scene = scenes.patch_room()
    for mode in ("ScatterBackup", "GatherOnly", "MultiScatter"):
        props = {"temporalReuse": mode, "gatherOption": "Clamped", "maxSurfaceBounces": 4, "numTimePartitions": 2}
        gpu, cpu = _pair(props, scene)
        for f in range(3):
            print(f"frame {f} glossy {mode}", flush=True)
            cam = scenes.box_room_camera(f, W, H, aperture=0.03, focal_distance=3.5,
                                         shutter=1.0 / 30.0 if mode == "MultiScatter" else 0.0)
            gpu.execute(cam, W, H)
            cpu.execute(cam, W, H)
            _compare(gpu, cpu, W, H, ["history_reservoirs", "history_recon", "color"], f)
            assert gpu.pixelStats == cpu.pixelStats



@pytest.mark.parametrize("mode,K", [("ScatterOnly", 1), ("ScatterBackup", 1), ("MultiScatter", 3)])
def test_many_splats_into_few_cells(mode, K):
    """The camera jumps from inside the room to 400 units away, so the previous frame's
    hits reproject onto a few pixels: up to ~6800 splats per cell, past the 2048-entry LDS
    chunk of the long-segment sort and through its merge passes.  Splat lists, sort order and
    the frame stay bit-exact with the oracle, and the lists keep their defining properties."""
    from test_gpu_batch import _splat_lists_consistent
    W, H = 192, 160
    scene = scenes.box_room(seed=0)
    gpu, cpu = _pair({"temporalReuse": mode, "numTimePartitions": K, "maxSurfaceBounces": 3}, scene)
    shutter = 1.0 / 30.0 if K > 1 else 0.0
    cams = [scenes.make_camera((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), 40.0, W / H, shutter=shutter),
            scenes.make_camera((0.0, 0.0, 400.0), (0.0, 0.0, 0.0), 40.0, W / H, shutter=shutter)]
    for f, cam in enumerate(cams):
        print(f"frame {f} many splats {mode}", flush=True)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
    cnt = cpu.read_buffer("cell_count", W, H, K).view(np.uint32)
    assert cnt.max() > 2048, cnt.max()
    for name in ("splat_dst", "cell_count", "cell_offset"):
        assert np.array_equal(gpu.read_buffer(name, W, H, K), cpu.read_buffer(name, W, H, K)), name
    total = int(cnt.sum())
    gs = gpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total]
    cs = cpu.read_buffer("sorted", W, H, K).view(np.uint32)[:total]
    assert np.array_equal(gs, cs)
    if K == 1:
        _splat_lists_consistent(gpu, W, H)
    _compare(gpu, cpu, W, H, ["temporal_reservoirs", "temporal_recon", "history_reservoirs", "color"], 1)
    assert gpu.pixelStats == cpu.pixelStats
