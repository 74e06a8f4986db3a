"""Imported scenes on the GPU: the Cornell box read from its pbrt-v4, Mitsuba 3 and
.pyscene files (tests/golden/scenes) and rendered through the pass with the file's own
camera, bit for bit against the oracle on the same imported arrays, in ScatterBackup and
GatherOnly; and the three files' renders agree with each other in the mean under one
diffuse model (they describe the same box, triangulated differently)."""
import os

import numpy as np
import pytest
import torch

import oracle_harness as oh
import rs_amd
from rs_amd import importers

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def _compare(gpu, cpu, W, H, names, frame):
    for name in names:
        g = gpu.read_buffer(name, W, H)
        c = cpu.read_buffer(name, W, H)
        bpp = rs_amd.abi.BUF_BYTES_PER_PIXEL[rs_amd.abi.BUF[name]]
        mm = oh.first_mismatch(g, c, W, bpp)
        assert mm is None, f"frame {frame} {name}: first mismatch (pixel, word, xy, gpu, cpu, #pixels) = {mm}"


@pytest.mark.parametrize("fmt", ["pbrt", "xml", "pyscene"])
@pytest.mark.parametrize("mode", ["ScatterBackup", "GatherOnly"])
def test_imported_cornell_bit_exact(fmt, mode):
    W, H = 64, 64
    s = importers.import_scene(os.path.join(SCENES, "cornell." + fmt))
    props = {"temporalReuse": mode, "maxSurfaceBounces": 4, "captureStages": 1}
    gpu = rs_amd.ReservoirSplatting(props)
    s.apply(gpu)
    cpu = oh.OraclePass(gpu._opts)
    cpu.setScene(s)
    cam = s.camera(W, H)
    for f in range(3):
        print(f"frame {f} imported cornell.{fmt} {mode}", flush=True)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, ["vbuffer", "initial_reservoirs", "initial_recon", "history_reservoirs",
                                  "history_recon", "color"], f)
        assert gpu.pixelStats == cpu.pixelStats


def test_three_formats_render_alike():
    W, H, n = 64, 64, 256
    means = {}
    for fmt in ("pbrt", "xml", "pyscene"):
        s = importers.import_scene(os.path.join(SCENES, "cornell." + fmt))
        # the Mitsuba import selects the Lambertian diffuse of its BSDFs; compare geometry and
        # lighting under one diffuse model
        s.diffuse_model = rs_amd.abi.DIFFUSE_FROSTBITE
        gpu = rs_amd.ReservoirSplatting({"temporalReuse": "ScatterBackup", "maxSurfaceBounces": 4})
        s.apply(gpu)
        acc = rs_amd.AccumulatePass({"precisionMode": "Double"})
        color = torch.empty((H, W, 4), device="cuda")
        out = torch.empty((H, W, 4), device="cuda")
        cam = s.camera(W, H)
        print(f"accumulating cornell.{fmt}", flush=True)
        for _ in range(n):
            gpu.execute(cam, W, H, color=color)
            acc.execute(color, out)
        torch.cuda.synchronize()
        means[fmt] = out[..., :3].double().mean(dim=(0, 1)).cpu().numpy()
    print(means, flush=True)
    for fmt in ("xml", "pyscene"):
        assert np.allclose(means[fmt], means["pbrt"], rtol=0.02), (fmt, means)
