"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle sanitize`
builds the oracle with -fsanitize=address,undefined and runs a driver through every temporal
mode, gather mechanism and option branch, the frozen scene, instance motion, a caller
VBuffer, the environment map, per-pixel stats, the FP64 replay and the accumulator."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out, out[-4000:]
    assert "AddressSanitizer" not in out and "LeakSanitizer" not in out, out[-4000:]
    assert "0 failures" in out, out[-2000:]
