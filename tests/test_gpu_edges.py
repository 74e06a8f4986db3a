"""Edge cases of the pass on the GPU, bit for bit against the oracle:
* scenes of 1, 2, 3 and 5 triangles: the BVH's single-leaf root (1 and 2 triangles) and
  the smallest clustered trees (3 and 5), with one emissive triangle;
* frames of 1x1, 2x3, 33x1 and 1x17 pixels: all but one lane of a 16x16 tile idle, and
  spatial neighbours that fall outside the frame;
* a resolution change between frames, which drops the history on both sides."""
import numpy as np
import pytest

import oracle_harness as oh
import rs_amd
from rs_amd import scenes

pytestmark = pytest.mark.gpu

BUFFERS = ["mvec", "vbuffer", "initial_reservoirs", "initial_recon", "history_reservoirs", "history_recon", "color"]


def _compare(gpu, cpu, W, H, frame):
    for name in BUFFERS:
        g = gpu.read_buffer(name, W, H)
        c = cpu.read_buffer(name, W, H)
        bpp = rs_amd.abi.BUF_BYTES_PER_PIXEL[rs_amd.abi.BUF[name]]
        mm = oh.first_mismatch(g, c, W, bpp)
        assert mm is None, f"frame {frame} {W}x{H} {name}: first mismatch (pixel, word, xy, gpu, cpu, #pixels) = {mm}"
    assert gpu.pixelStats == cpu.pixelStats


def _run(scene, mode, sizes, camera):
    props = {"temporalReuse": mode, "maxSurfaceBounces": 4, "captureStages": 1}
    gpu = rs_amd.ReservoirSplatting(props)
    gpu.setScene(scene)
    cpu = oh.OraclePass(gpu._opts)
    cpu.setScene(scene)
    for f, (W, H) in enumerate(sizes):
        print(f"frame {f} {mode} {W}x{H} triangles {scene.num_triangles}", flush=True)
        cam = camera(f, W, H)
        gpu.execute(cam, W, H)
        cpu.execute(cam, W, H)
        _compare(gpu, cpu, W, H, f)


def _tiny_scene(n):
    """`n` triangles: an emissive triangle facing down at y = 1.5, then a floor quad at
    y = 0 and a back-wall quad at z = -2 as far as `n` reaches."""
    s = scenes.Scene()
    light = s.add_material((0.0, 0.0, 0.0), emissive=(6.0, 5.0, 4.0))
    floor = s.add_material((0.7, 0.6, 0.5), roughness=0.6)
    wall = s.add_material((0.3, 0.5, 0.7), roughness=0.3, metallic=0.2)
    s.add_mesh([[-0.4, 1.5, -0.4], [0.0, 1.5, 0.4], [0.4, 1.5, -0.4]], [[0.0, -1.0, 0.0]] * 3, [[0, 2, 1]], light)
    quads = [([[-2.0, 0.0, -2.0], [-2.0, 0.0, 2.0], [2.0, 0.0, 2.0], [2.0, 0.0, -2.0]], [0.0, 1.0, 0.0], floor),
             ([[-2.0, 0.0, -2.0], [2.0, 0.0, -2.0], [2.0, 2.5, -2.0], [-2.0, 2.5, -2.0]], [0.0, 0.0, 1.0], wall)]
    left = n - 1
    for P, nrm, mat in quads:
        if left == 0:
            break
        tris = [[0, 1, 2], [0, 2, 3]][:min(left, 2)]
        s.add_mesh(P, [nrm] * 4, tris, mat)
        left -= len(tris)
    assert left == 0
    s = s.finalize()
    assert s.num_triangles == n
    return s


def _tiny_camera(frame, W, H):
    return scenes.make_camera((0.1 * frame, 1.0, 3.0), (0.0, 0.6, 0.0), 50.0, W / H)


@pytest.mark.parametrize("mode", ["ScatterBackup", "GatherOnly"])
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_tiny_scenes_bit_exact(n, mode):
    _run(_tiny_scene(n), mode, [(23, 17)] * 3, _tiny_camera)


@pytest.mark.parametrize("mode", ["ScatterBackup", "GatherOnly"])
@pytest.mark.parametrize("W,H", [(1, 1), (2, 3), (33, 1), (1, 17)])
def test_tiny_frames_bit_exact(W, H, mode):
    _run(scenes.box_room(seed=0), mode, [(W, H)] * 3, scenes.box_room_camera)


def test_resize_between_frames_bit_exact():
    sizes = [(40, 24), (40, 24), (24, 40), (24, 40), (40, 24)]
    _run(scenes.box_room(seed=0), "ScatterBackup", sizes, scenes.box_room_camera)
