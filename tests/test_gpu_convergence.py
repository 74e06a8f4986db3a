"""Unbiasedness on the GPU: with temporal and spatial reuse the accumulated image converges
to the converged accumulation of the plain path tracer (relMSE as defined by ImageCompare,
Source/Tools/ImageCompare/ImageCompare.cpp:202-211).  Static camera on the box room; the
reference accumulation uses 8x as many frames as the reuse accumulation.  Both are
accumulated by the product's AccumulatePass in Double precision (the reference's converged
reference is the same pass after the path)."""
import numpy as np
import pytest
import torch

import rs_amd
from rs_amd import abi, metrics, scenes

ROBUST_COLOUR = abi.QUIRKS_REFERENCE & ~abi.QUIRK["robustGrey"]

pytestmark = pytest.mark.gpu


def _accumulate(props, frames, W, H, shutter=0.0):
    rs = rs_amd.ReservoirSplatting(props)
    rs.setScene(scenes.box_room(seed=0))
    acc = rs_amd.AccumulatePass({"precisionMode": "Double"})
    cam = scenes.box_room_camera(0, W, H, shutter=shutter)
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    for _ in range(frames):
        rs.execute(cam, W, H, color=color)
        acc.execute(color, out)
    torch.cuda.synchronize()
    assert acc.frameCount == frames
    return out.double().cpu().numpy()


# The Robust gather converges in colour only with its grey-integrand quirk corrected
# (ROBUST_COLOUR clears RS_QUIRK_ROBUST_GREY from the reference's mask).
@pytest.mark.parametrize("mode,gather,shutter,quirks", [("ScatterBackup", "Clamped", 0.0, abi.QUIRKS_REFERENCE),
                                                        ("ScatterOnly", "Clamped", 0.0, abi.QUIRKS_REFERENCE),
                                                        ("GatherOnly", "Fast", 0.0, abi.QUIRKS_REFERENCE),
                                                        ("MultiScatter", "Clamped", 1.0 / 30.0, abi.QUIRKS_REFERENCE),
                                                        ("ScatterBackup", "Robust", 0.0, ROBUST_COLOUR)])
def test_reuse_converges_to_path_traced_accumulation(mode, gather, shutter, quirks):
    W, H = 64, 64
    print(f"convergence {mode} {gather} quirks {quirks}", flush=True)
    base = {"maxSurfaceBounces": 4, "samplesPerPixel": 1}
    ref = _accumulate({**base, "enableTemporalResampling": False, "enableSpatialResampling": False}, 16384, W, H)
    r = _accumulate({**base, "temporalReuse": mode, "gatherOption": gather, "numTimePartitions": 2,
                     "quirkMask": quirks}, 2048, W, H, shutter)
    rel = metrics.rel_mse(r, ref)
    print(f"relMSE {rel:.5f} mean ratio {r[..., :3].mean() / ref[..., :3].mean():.5f}", flush=True)
    assert rel < 0.01, rel
    assert abs(r[..., :3].mean() / ref[..., :3].mean() - 1) < 0.01


def _replay_accumulate(scene, props, cams, samples, W, H, replay_accumulation=True):
    """Accumulate the last frame of the camera sequence over `samples` replays of the whole
    sequence (CameraReplay, "Replay for Accumulation"): every replay restarts the history
    (Scene::shouldReset on its first frame) with fresh seeds, and the AccumulatePass adds only
    the frames the pass marks with _shouldAccumulate.  Without replay_accumulation the scene
    freezes on the last frame instead, and `samples` re-renders of it are accumulated (the
    path-traced reference, which keeps no history)."""
    rs = rs_amd.ReservoirSplatting(props)
    rs.setScene(scene)
    acc = rs_amd.AccumulatePass({"precisionMode": "Double"})
    replay = rs_amd.CameraReplay(cams, accumulation_frame=len(cams) - 1, replay_accumulation=replay_accumulation)
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    while acc.frameCount < samples:
        flags = replay.execute(rs, W, H, color=color)
        acc.execute(color, out, shouldAccumulate=flags["_shouldAccumulate"])
    torch.cuda.synchronize()
    return out.double().cpu().numpy()


# GGX + Frostbite patch room (roughness U[0.2, 0.9], so both sides of the 0.25 specular
# threshold): every temporal mode, gather mechanism and backup MIS under camera motion.
GLOSSY = [("ScatterBackup", "Clamped", "Balance", 0.0, 0.0, abi.QUIRKS_REFERENCE), ("ScatterBackup", "Fast", "Pairwise", 0.0, 0.0, abi.QUIRKS_REFERENCE),
          ("ScatterOnly", "Clamped", "Balance", 0.0, 0.0, abi.QUIRKS_REFERENCE), ("GatherOnly", "Fast", "Balance", 0.0, 0.0, abi.QUIRKS_REFERENCE),
          ("GatherOnly", "Robust", "Balance", 0.0, 0.0, ROBUST_COLOUR), ("MultiScatter", "Clamped", "Balance", 1.0 / 30.0, 0.0, abi.QUIRKS_REFERENCE),
          ("ScatterBackup", "Clamped", "Balance", 0.0, 0.03, abi.QUIRKS_REFERENCE), ("GatherOnly", "Clamped", "Balance", 0.0, 0.03, abi.QUIRKS_REFERENCE)]


@pytest.mark.parametrize("mode,gather,backup,shutter,aperture,quirks", GLOSSY)
def test_glossy_reuse_converges_under_motion(mode, gather, backup, shutter, aperture, quirks):
    W, H = 64, 64
    print(f"glossy motion {mode} {gather} {backup} shutter {shutter} aperture {aperture}", flush=True)
    scene = scenes.patch_room()
    cams = [scenes.box_room_camera(f, W, H, shutter=shutter, aperture=aperture, focal_distance=3.5, speed=0.04)
            for f in range(4)]
    base = {"maxSurfaceBounces": 4, "samplesPerPixel": 1, "numTimePartitions": 2, "quirkMask": quirks}
    ref = _replay_accumulate(scene, {**base, "enableTemporalResampling": False, "enableSpatialResampling": False},
                             cams, 32768, W, H, replay_accumulation=False)
    r = _replay_accumulate(scene, {**base, "temporalReuse": mode, "gatherOption": gather,
                                   "scatterBackupMISOption": backup}, cams, 2048, W, H)
    rel = metrics.rel_mse(r, ref)
    print(f"relMSE {rel:.5f} mean ratio {r[..., :3].mean() / ref[..., :3].mean():.5f}", flush=True)
    assert rel < 0.01, rel
    assert abs(r[..., :3].mean() / ref[..., :3].mean() - 1) < 0.01
