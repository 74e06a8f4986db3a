"""The frozen-replay adjoint on the GPU (torch.autograd.Function over rs_replay_forward /
rs_replay_backward) against the oracle's FP64 dual-number gradient and against central
finite differences (eps 1e-3) of the GPU replay itself, on the patch room with 64 patch
materials (base colour and roughness), L2 loss against a target image; also at the
differentiable workload's full size, 512 x 512."""
import numpy as np
import pytest
import torch

import oracle_harness as oh
import rs_amd
from rs_amd import scenes
from rs_amd.adjoint import replay

pytestmark = pytest.mark.gpu

def _setup(patches, W, H, sky=False):
    scene = scenes.patch_room(patches_per_wall=patches)
    props = {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "maxSurfaceBounces": 4}
    gpu = rs_amd.ReservoirSplatting(props)
    gpu.setScene(scene)
    cpu = oh.OraclePass(gpu._opts)
    cpu.setScene(scene)
    if sky:
        # light through the open front: paths whose light vertex is the distant environment
        env = scenes.sky_env_map(64, 32)
        gpu.setEnvMap(env, intensity=0.5, rotation=(10.0, 180.0, 0.0))
        cpu.setEnvMap(env, intensity=0.5, rotation=(10.0, 180.0, 0.0))
    for f in range(3):
        cam = scenes.box_room_camera(0, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
    assert np.array_equal(gpu.read_buffer("color", W, H), cpu.read_buffer("color", W, H))
    gpu.freeze()
    theta0 = gpu.materialParams()
    return gpu, cpu, theta0


def _perturb(theta0, seed, scale_a, scale_r):
    rng = np.random.default_rng(seed)
    t = theta0.astype(np.float64).copy()
    t[:, :3] *= rng.uniform(1 - scale_a, 1 + scale_a, (t.shape[0], 3))
    t[:, 3] = np.clip(t[:, 3] + rng.uniform(-scale_r, scale_r, t.shape[0]), 0.15, 1.0)
    return t.astype(np.float32)


# 68 materials (block sums in LDS), 404 (wave sums); and under an environment map
@pytest.mark.parametrize("patches,W,H,sky", [(16, 96, 80, False), (100, 96, 80, False), (16, 512, 512, False),
                                             (16, 96, 80, True)])
def test_adjoint_matches_fp64_and_finite_differences(patches, W, H, sky):
    print(f"adjoint: setup, {patches} patches per wall, {W}x{H}, sky {sky}", flush=True)
    gpu, cpu, theta0 = _setup(patches, W, H, sky)
    dev = torch.device("cuda")
    # the replay at the scene's parameters is the frozen frame
    r0 = replay(gpu, torch.from_numpy(theta0).to(dev), H, W)
    assert torch.equal(r0, torch.from_numpy(gpu.read_buffer("color", W, H).view(np.float32).reshape(H, W, 4)).to(dev))
    target_theta = _perturb(theta0, 5, 0.2, 0.1)
    target = replay(gpu, torch.from_numpy(target_theta).to(dev), H, W).detach()
    theta1 = _perturb(theta0, 6, 0.1, 0.05)
    p1 = torch.from_numpy(theta1).to(dev).requires_grad_(True)
    loss = ((replay(gpu, p1, H, W).double() - target.double()) ** 2).sum()
    loss.backward()
    g_auto = p1.grad.double().cpu().numpy()
    # raw adjoint in double
    g_gpu = torch.empty(theta1.shape, dtype=torch.float64, device=dev)
    R = replay(gpu, torch.from_numpy(theta1).to(dev), H, W)
    gpu.replay_backward(torch.from_numpy(theta1).to(dev), (2.0 * (R - target)).contiguous(), g_gpu)
    g_gpu = g_gpu.cpu().numpy()
    assert np.allclose(g_auto, g_gpu, rtol=1e-6, atol=1e-7)
    print("adjoint: fp64", flush=True)
    _, L_cpu, g_cpu = cpu.replay(theta1.astype(np.float64), W, H, target.cpu().numpy())
    assert abs(loss.item() - L_cpu) <= 1e-4 * L_cpu
    rel = np.linalg.norm(g_gpu - g_cpu) / np.linalg.norm(g_cpu)
    print(f"relative error vs FP64: {rel:.2e}", flush=True)
    assert rel < 1e-3, rel
    big = np.abs(g_cpu) > 1e-2 * np.abs(g_cpu).max()
    assert np.allclose(g_gpu[big], g_cpu[big], rtol=1e-3)
    print("adjoint: finite differences", flush=True)
    eps = 1e-3
    flat = theta1.ravel()
    for k in np.argsort(-np.abs(g_cpu).ravel())[:16]:
        tp, tm = flat.copy(), flat.copy()
        tp[k] += eps
        tm[k] -= eps
        Lp = ((replay(gpu, torch.from_numpy(tp.reshape(theta1.shape)).to(dev), H, W).double() - target.double()) ** 2).sum()
        Lm = ((replay(gpu, torch.from_numpy(tm.reshape(theta1.shape)).to(dev), H, W).double() - target.double()) ** 2).sum()
        fd = (Lp - Lm).item() / ((tp[k] - tm[k]).astype(np.float64))
        assert abs(fd - g_gpu.ravel()[k]) <= 1e-3 * abs(g_gpu.ravel()[k]), (int(k), fd, g_gpu.ravel()[k])
