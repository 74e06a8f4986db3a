"""Pins of the AccumulatePass oracle (oracle/accumulate.cpp) by what the mathematics and
the reference's host logic fix, independently of its formulas:

* Double mode is the mean of the frames, to float32 rounding (exact mean from math.fsum);
* Single mode is the mean within the float32 running-sum error bound;
* SingleCompensated recovers contributions that a plain float32 sum loses (Kahan), so a
  dropped correction term fails;
* the moving average (maxFrameCount > 0) is the running mean up to the limit and then an
  exponential blend with weight 1/(limit+1): a step input decays as (M/(M+1))^j;
* overflow Stop leaves the output untouched, Reset restarts at the current frame;
* view accumulation shows the mean without adding a frame; a disabled pass copies; a new
  frame size restarts (AccumulatePass.cpp:130-198, 270-290).
"""
import math

import numpy as np
import pytest

import oracle_harness as oh

SINGLE, COMP, DOUBLE = 0, 1, 2
STOP, RESET, EMA = 0, 1, 2


def _frames(n, H=5, W=7, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray((rng.random((H, W, 4)) * scale).astype(np.float32)) for _ in range(n)]


def _run(acc, frames, **kw):
    outs = []
    for f in frames:
        d = np.full_like(f, np.nan)
        acc.execute(f, d, **kw)
        outs.append(d)
    return outs


def _exact_mean(frames):
    stack = np.stack([f.astype(np.float64) for f in frames])
    flat = stack.reshape(len(frames), -1)
    return np.array([math.fsum(flat[:, j]) / len(frames) for j in range(flat.shape[1])]).reshape(frames[0].shape)


def test_double_is_the_mean_to_float_rounding():
    frames = _frames(64, seed=1)
    outs = _run(oh.OracleAccum(precision=DOUBLE), frames)
    for n in (1, 2, 7, 64):
        want = _exact_mean(frames[:n]).astype(np.float32)
        got = outs[n - 1]
        ulp = np.spacing(np.abs(want))
        assert np.all(np.abs(got.astype(np.float64) - want) <= ulp), n


def test_single_is_the_mean_within_the_sum_bound():
    frames = _frames(100, seed=2)
    outs = _run(oh.OracleAccum(precision=SINGLE), frames)
    eps = np.finfo(np.float32).eps
    for n in (1, 10, 100):
        exact = _exact_mean(frames[:n])
        bound = (n + 2) * eps * np.max([np.abs(f) for f in frames[:n]], axis=0)
        assert np.all(np.abs(outs[n - 1] - exact) <= bound + 1e-12), n


def test_compensated_keeps_what_a_float_sum_loses():
    H, W = 2, 3
    big = np.full((H, W, 4), 1.0e8, np.float32)
    ones = np.ones((H, W, 4), np.float32)
    frames = [big] + [ones] * 999
    exact = (1.0e8 + 999.0) / 1000.0
    comp = _run(oh.OracleAccum(precision=COMP), frames)[-1]
    single = _run(oh.OracleAccum(precision=SINGLE), frames)[-1]
    assert np.all(np.abs(comp - exact) <= 2e-7 * exact)
    # 1e8 + 1 rounds back to 1e8 in float32: the plain sum loses every one
    assert np.all(np.abs(single - 1.0e5) <= 1e-2)


@pytest.mark.parametrize("precision", [SINGLE, DOUBLE])
def test_moving_average_mean_then_exponential_blend(precision):
    M, H, W = 8, 3, 4
    c0 = np.full((H, W, 4), 2.0, np.float32)
    c1 = np.full((H, W, 4), 10.0, np.float32)
    acc = oh.OracleAccum(precision=precision, maxFrameCount=M, overflowMode=EMA)
    outs = _run(acc, [c0] * M)
    assert np.allclose(outs[-1], 2.0, rtol=1e-6)
    assert acc.frameCount == M
    outs = _run(acc, [c1] * 20)
    for j, o in enumerate(outs, start=1):
        want = 10.0 + (2.0 - 10.0) * (M / (M + 1.0)) ** j
        assert np.allclose(o, want, rtol=2e-5), j
    assert acc.frameCount == M


def test_overflow_stop_and_reset():
    M = 5
    frames = _frames(M + 3, seed=3)
    acc = oh.OracleAccum(precision=SINGLE, maxFrameCount=M, overflowMode=STOP)
    outs = _run(acc, frames)
    for o in outs[M:]:
        assert np.all(np.isnan(o)), "Stop must not write the output"
    assert acc.frameCount == M
    acc = oh.OracleAccum(precision=SINGLE, maxFrameCount=M, overflowMode=RESET)
    outs = _run(acc, frames)
    assert np.array_equal(outs[M], frames[M])  # restarted: the mean of one frame
    assert acc.frameCount == 3  # the three frames since the restart


def test_view_disabled_and_resize():
    frames = _frames(4, seed=4)
    acc = oh.OracleAccum(precision=DOUBLE)
    outs = _run(acc, frames[:3])
    view = np.full_like(frames[3], np.nan)
    acc.execute(frames[3], view, shouldAccumulate=False, viewAccumulation=True)
    assert np.array_equal(view, outs[2]) or np.allclose(view, outs[2], rtol=1e-7)
    assert acc.frameCount == 3
    plain = np.full_like(frames[3], np.nan)
    acc.execute(frames[3], plain, shouldAccumulate=False)
    assert np.array_equal(plain, frames[3])
    other = _frames(1, H=6, W=7, seed=5)[0]
    out = np.full_like(other, np.nan)
    acc.execute(other, out)
    assert np.array_equal(out, other) and acc.frameCount == 1
    off = oh.OracleAccum(enabled=0)
    d = np.zeros_like(frames[0])
    off.execute(frames[0], d)
    assert np.array_equal(d, frames[0])
