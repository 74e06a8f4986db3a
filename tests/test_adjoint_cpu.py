"""The FP64 replay gradient of the oracle (forward-mode dual numbers) against central finite
differences of its own double-precision replay, on the patch room (64 patch materials with
GGX and Frostbite lobes)."""
import ctypes as C

import numpy as np

import oracle_harness as oh
import rs_amd
from rs_amd import abi, scenes


def _frozen_oracle(W=40, H=32, frames=2):
    o = abi.Options()
    rs_amd.load_library().rs_default_options(C.byref(o))
    o.maxSurfaceBounces = o.maxDiffuseBounces = o.maxSpecularBounces = o.maxTransmissionBounces = 4
    o.temporalReuse = abi.TEMPORAL["ScatterBackup"]
    o.gatherMechanism = abi.GATHER["Clamped"]
    scene = scenes.patch_room()
    p = oh.OraclePass(o)
    p.setScene(scene)
    for f in range(frames):
        p.execute(scenes.box_room_camera(0, W, H), W, H)
    theta0 = np.asarray([list(m[0]) + [m[1]] for m in scene.materials], np.float32).astype(np.float64)
    return p, scene, theta0


def test_replay_reproduces_frame_at_scene_parameters():
    W, H = 40, 32
    p, scene, theta0 = _frozen_oracle(W, H)
    img, _, _ = p.replay(theta0, W, H)
    assert np.array_equal(img, p.color[..., :3].astype(np.float64))


def test_dual_gradient_matches_central_differences():
    W, H = 40, 32
    p, scene, theta0 = _frozen_oracle(W, H)
    rng = np.random.default_rng(5)
    target_theta = theta0.copy()
    target_theta[:, :3] *= rng.uniform(0.8, 1.2, (theta0.shape[0], 3))
    target_theta[:, 3] = np.clip(theta0[:, 3] + rng.uniform(-0.1, 0.1, theta0.shape[0]), 0.15, 1.0)
    tgt, _, _ = p.replay(target_theta, W, H)
    target = np.concatenate([tgt, np.ones((H, W, 1))], -1).astype(np.float32)
    theta1 = theta0.copy()
    theta1[:, :3] *= rng.uniform(0.9, 1.1, (theta0.shape[0], 3))
    theta1[:, 3] = np.clip(theta0[:, 3] + rng.uniform(-0.05, 0.05, theta0.shape[0]), 0.15, 1.0)
    _, L, grad = p.replay(theta1, W, H, target)
    assert L > 0 and np.abs(grad).max() > 0
    order = np.argsort(-np.abs(grad).ravel())[:24]
    eps = 1e-5
    for k in order:
        tp, tm = theta1.copy().ravel(), theta1.copy().ravel()
        tp[k] += eps
        tm[k] -= eps
        _, Lp, _ = p.replay(tp.reshape(theta1.shape), W, H, target)
        _, Lm, _ = p.replay(tm.reshape(theta1.shape), W, H, target)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - grad.ravel()[k]) <= 1e-5 * abs(grad.ravel()[k]) + 1e-9, (k, fd, grad.ravel()[k])
    # roughness gradients exist (GGX and Frostbite terms are exercised)
    assert np.abs(grad[:, 3]).max() > 0


def test_differentiable_room_has_64_materials_and_seed_independent_geometry():
    """The differentiable workload's room: exactly 64 materials, every one used by some
    triangle, one of them emissive; the seeds change the parameters, never the geometry."""
    a, b = scenes.differentiable_room(5), scenes.differentiable_room(6)
    assert len(a.materials) == len(b.materials) == 64
    tm = np.asarray(a.tri_material)
    assert set(np.unique(tm).tolist()) == set(range(64))
    assert np.array_equal(tm, np.asarray(b.tri_material))
    assert np.array_equal(np.asarray(a.positions), np.asarray(b.positions))
    ta = np.asarray([list(m[0]) + [m[1]] for m in a.materials])
    tb = np.asarray([list(m[0]) + [m[1]] for m in b.materials])
    assert not np.allclose(ta, tb)
    assert sum(1 for m in a.material_array if max(m.emissive) > 0) == 1
