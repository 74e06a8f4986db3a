"""AccumulatePass on the GPU (rs_accum_* through rs_amd.AccumulatePass) against the oracle
(oracle/accumulate.cpp), bit for bit: every precision mode, the moving average with each
overflow mode, view accumulation, frames that are not accumulated, the disabled copy and a
frame-size change; and at 1920x1080, the Double mean of 24 frames against the exact mean."""
import math

import numpy as np
import pytest
import torch

import oracle_harness as oh
import rs_amd

pytestmark = pytest.mark.gpu

PRECISIONS = {"Single": 0, "SingleCompensated": 1, "Double": 2}
OVERFLOWS = {"Stop": 0, "Reset": 1, "EMA": 2}


def _frames(n, H, W, seed):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray((rng.random((H, W, 4)) * 4.0).astype(np.float32)) for _ in range(n)]


def _step(gpu, cpu, src, keep, **kw):
    """One execute on both; `keep` holds each side's previous output (Stop leaves it)."""
    dev = torch.from_numpy(src).cuda()
    gout = keep[0]
    gpu.execute(dev, gout, **kw)
    cout = keep[1]
    cpu.execute(src, cout, **kw)
    torch.cuda.synchronize()
    g = gout.cpu().numpy()
    assert np.array_equal(g.view(np.uint32), cout.view(np.uint32)), (kw, gpu.getProperties(), gpu.frameCount)
    assert gpu.frameCount == cpu.frameCount


@pytest.mark.parametrize("precision", list(PRECISIONS))
@pytest.mark.parametrize("maxFrames,overflow", [(0, "Stop"), (6, "Stop"), (6, "Reset"), (6, "EMA")])
def test_accumulate_bit_exact(precision, maxFrames, overflow):
    H, W = 61, 97
    print(f"accumulate {precision} max {maxFrames} {overflow}", flush=True)
    props = {"precisionMode": precision, "maxFrameCount": maxFrames, "overflowMode": overflow}
    gpu = rs_amd.AccumulatePass(props)
    cpu = oh.OracleAccum(1, PRECISIONS[precision], maxFrames, OVERFLOWS[overflow])
    keep = (torch.full((H, W, 4), float("nan"), device="cuda"), np.full((H, W, 4), np.nan, np.float32))
    frames = _frames(16, H, W, seed=7)
    for k, f in enumerate(frames):
        _step(gpu, cpu, f, keep)
        if k == 9:
            _step(gpu, cpu, f, keep, shouldAccumulate=False, viewAccumulation=True)
            _step(gpu, cpu, f, keep, shouldAccumulate=False)


def test_accumulate_resize_disable_reset():
    gpu = rs_amd.AccumulatePass({"precisionMode": "Double"})
    cpu = oh.OracleAccum(1, 2, 0, 0)
    for H, W in ((40, 50), (40, 50), (33, 50), (33, 50)):
        f = _frames(1, H, W, seed=H + W)[0]
        keep = (torch.full((H, W, 4), float("nan"), device="cuda"), np.full((H, W, 4), np.nan, np.float32))
        _step(gpu, cpu, f, keep)
    gpu.reset()
    cpu.reset()
    f = _frames(1, 33, 50, seed=1)[0]
    keep = (torch.zeros((33, 50, 4), device="cuda"), np.zeros((33, 50, 4), np.float32))
    _step(gpu, cpu, f, keep)
    gpu.enabled = False
    cpu.set_options(enabled=0, precision=2)
    _step(gpu, cpu, f, keep)
    assert np.array_equal(keep[1], f)


def test_accumulate_1080p_double_mean():
    H, W, n = 1080, 1920, 24
    gen = torch.Generator(device="cuda").manual_seed(3)
    gpu = rs_amd.AccumulatePass({"precisionMode": "Double"})
    out = torch.empty((H, W, 4), device="cuda")
    total = torch.zeros((H, W, 4), dtype=torch.float64, device="cuda")
    for _ in range(n):
        f = torch.rand((H, W, 4), generator=gen, device="cuda")
        gpu.execute(f, out)
        total += f.double()
    torch.cuda.synchronize()
    mean = (total / n).float()
    # the running double sum and the product's float conversion agree with the mean of the
    # same frames to one float32 ulp
    ulp = torch.abs(mean).nextafter(torch.tensor(math.inf, device="cuda")) - torch.abs(mean)
    assert torch.all(torch.abs(out - mean) <= ulp)
    assert gpu.frameCount == n
