"""The C++ host mirror of the pass on the GPU: Cornell-box frames (ScatterBackup/Clamped,
GatherOnly/Robust, MultiScatter with a shutter) rendered through rs_amd::ReservoirSplatting
equal the oracle's bit for bit, and reset() restarts the history (tests/cpp/host_mirror_test)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_mirror_test")


def test_host_mirror_gpu_cases():
    assert os.path.exists(BIN), f"{BIN} is not built (run __graft_entry__.build())"
    print("host mirror gpu cases", flush=True)
    r = subprocess.run([BIN, "gpu", ROOT], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 4
