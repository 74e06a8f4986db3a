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


def _splat_lists_consistent(p, w, h):
    """Defining properties of a frame's splat lists: the offsets are the exclusive scan of
    the counts, every splatted source appears exactly once, in its destination cell, and the
    sources inside a cell ascend."""
    dst = p.read_buffer("splat_dst", w, h).view(np.uint32)
    cnt = p.read_buffer("cell_count", w, h).view(np.uint32)
    off = p.read_buffer("cell_offset", w, h).view(np.uint32)
    srt = p.read_buffer("sorted", w, h).view(np.uint32)
    live = dst != 0xFFFFFFFF
    assert int(cnt.sum()) == int(live.sum())
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(cnt, dtype=np.uint64)[:-1]]).astype(np.uint32))
    total = int(cnt.sum())
    entries = srt[:total]
    cells = np.repeat(np.arange(cnt.size, dtype=np.uint32), cnt)
    assert np.array_equal(dst[entries], cells)
    assert np.array_equal(np.sort(entries), np.flatnonzero(live).astype(np.uint32))
    same = cells[1:] == cells[:-1]
    assert np.all(entries[1:][same] > entries[:-1][same])


def test_batch_workload_16_scenes():
    """The 16-scene batched workload (resolutions 320x240 .. 1920x1080, 1e4 .. 2e6 triangles,
    ScatterBackup / Clamped, max depth 6): batched frames equal the individual frames bit for
    bit on every scene, the smallest scenes equal the oracle bit for bit, and every scene's
    splat lists satisfy their defining properties."""
    import oracle_harness as oh
    work = scenes.batch_workload()
    props = {"samplesPerPixel": 1, "maxSurfaceBounces": 6, "temporalReuse": "ScatterBackup", "gatherOption": "Clamped"}

    def make():
        ps = []
        for sc, w, h in work:
            p = rs_amd.ReservoirSplatting(props)
            p.setScene(sc)
            ps.append(p)
        return ps

    batched, single = make(), make()
    smallest = sorted(range(len(work)), key=lambda i: (work[i][1] * work[i][2], work[i][0].num_triangles))[:2]
    oracles = {}
    for i in smallest:
        oracles[i] = oh.OraclePass(single[i]._opts)
        oracles[i].setScene(work[i][0])
    sizes = [(w, h) for _, w, h in work]
    batch = Batch(batched)
    for f in range(3):
        print(f"batch workload frame {f}", flush=True)
        cams = [scenes.orbit_camera(f, w, h) for w, h in sizes]
        batch.render(cams, sizes)
        for p, cam, (w, h) in zip(single, cams, sizes):
            p.execute(cam, w, h)
        for i in smallest:
            oracles[i].execute(cams[i], *sizes[i])
        for i, (p, q, (w, h)) in enumerate(zip(batched, single, sizes)):
            for name in ("history_reservoirs", "history_recon", "color"):
                assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), (f, i, name)
            assert p.pixelStats == q.pixelStats, (f, i)
            if f > 0:
                _splat_lists_consistent(p, w, h)
        for i in smallest:
            w, h = sizes[i]
            for name in ("history_reservoirs", "history_recon", "color"):
                assert np.array_equal(single[i].read_buffer(name, w, h), oracles[i].read_buffer(name, w, h)), (f, i, name)
            assert single[i].pixelStats == oracles[i].pixelStats
    batch.close()
