"""The C++ host mirror of the pass (reservoir-splatting_amd/host/ReservoirSplatting.h): its
property surface against the reference's parseProperties / validateOptions / getProperties,
run as the C++ test program tests/cpp/host_mirror_test (built by __graft_entry__.build())."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_mirror_test")


def test_host_mirror_cpu_cases():
    assert os.path.exists(BIN), f"{BIN} is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, "cpu", ROOT], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 9
