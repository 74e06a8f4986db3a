"""Batched frames (rs_batch_*): several contexts with different scenes, sizes and temporal
modes rendered per call as one hipGraph launch give exactly the buffers of rendering each
context on its own, for every queue count / tile width tried, and the autotuner returns a
valid configuration."""
import numpy as np
import pytest

import rs_amd
from rs_amd import scenes
from rs_amd.batch import Batch

pytestmark = pytest.mark.gpu

SETUPS = [({"temporalReuse": "ScatterBackup", "gatherOption": "Clamped"}, 96, 64, 0),
          ({"temporalReuse": "GatherOnly", "gatherOption": "Fast"}, 64, 48, 1),
          ({"temporalReuse": "ScatterOnly"}, 40, 72, 2)]


def _passes():
    out = []
    for props, w, h, seed in SETUPS:
        p = rs_amd.ReservoirSplatting({**props, "maxSurfaceBounces": 4})
        p.setScene(scenes.box_room(seed=seed))
        out.append(p)
    return out


@pytest.mark.parametrize("queues,tile", [(4, 16), (1, 8), (2, 64)])
def test_batch_matches_individual_frames(queues, tile):
    print(f"batch queues={queues} tile={tile}", flush=True)
    batched, single = _passes(), _passes()
    batch = Batch(batched)
    batch.configure(queues, tile)
    sizes = [(w, h) for _, w, h, _ in SETUPS]
    for f in range(4):
        cams = [scenes.box_room_camera(f, w, h) for w, h in sizes]
        batch.render(cams, sizes)
        for p, cam, (w, h) in zip(single, cams, sizes):
            p.execute(cam, w, h)
        for p, q, (w, h) in zip(batched, single, sizes):
            for name in ("history_reservoirs", "history_recon", "color"):
                assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), (f, name)
            assert p.pixelStats == q.pixelStats
    batch.close()


def test_batch_autotune():
    batched = _passes()
    batch = Batch(batched)
    sizes = [(w, h) for _, w, h, _ in SETUPS]
    cams = [scenes.box_room_camera(0, w, h) for w, h in sizes]
    q, t, ms = batch.autotune(cams, sizes, frames=1)
    assert q in (1, 2, 4, 8) and t in (8, 16, 32, 64) and ms > 0
    batch.render(cams, sizes)
    batch.close()
