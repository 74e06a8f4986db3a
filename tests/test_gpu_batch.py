"""Batched frames (rs_batch_*): several contexts with different scenes, sizes and temporal
modes rendered per call as one hipGraph launch give exactly the buffers of rendering each
context on its own, in the fused form (queues 0: one launch per stage for all contexts) and
the stream form (queues 1..8), for every tile width tried, and the autotuner returns a valid
configuration."""
import numpy as np
import pytest
import torch

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


def _configure(batch, form, count, tile):
    if form == "fused":
        batch.configure_fused(count, tile)
    else:
        batch.configure(count, tile)
    cfg = batch.config()
    assert cfg["form"] == form and cfg["tile_width"] == tile
    assert cfg["groups" if form == "fused" else "queues"] == count


@pytest.mark.parametrize("form,count,tile", [("fused", 1, 16), ("fused", 2, 8), ("fused", 3, 64), ("streams", 4, 16),
                                             ("streams", 1, 8), ("streams", 2, 64)])
def test_batch_matches_individual_frames(form, count, tile):
    print(f"batch {form} {count} tile={tile}", flush=True)
    batched, single = _passes(), _passes()
    batch = Batch(batched)
    _configure(batch, form, count, tile)
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


# Contexts whose options, instantiations and outputs all differ, batched together: every
# temporal mode and gather, lens and shutter, two spatial iterations of three neighbours,
# spatial reuse off, environment maps, per-pixel ray counts and path lengths, the caller's
# motion vectors, the stage capture, and odd frame sizes.
MIXED = [
    ({"temporalReuse": "ScatterBackup", "gatherOption": "Clamped"}, 96, 64, 0, False, {}, ()),
    ({"temporalReuse": "GatherOnly", "gatherOption": "Robust"}, 64, 48, 1, False, {"aperture": 0.05}, ()),
    ({"temporalReuse": "MultiScatter", "numTimePartitions": 3, "neighborCount": 2}, 40, 72, 2, False,
     {"shutter": 1.0 / 24.0, "speed": 0.05}, ()),
    ({"temporalReuse": "ScatterOnly", "neighborCount": 3, "numSpatialIterations": 2}, 48, 40, 3, True, {}, ()),
    ({"temporalReuse": "ScatterBackup", "gatherOption": "Robust", "scatterBackupMISOption": "Pairwise"}, 33, 17, 4,
     True, {}, ("rayCount", "pathLength")),
    ({"temporalReuse": "ScatterOnly", "captureStages": 1, "enableSpatialResampling": 0}, 24, 24, 5, False, {},
     ("mvec", "rayCount")),
    ({"temporalReuse": "GatherOnly", "gatherOption": "Clamped", "neighborCount": 0}, 17, 33, 6, True,
     {"aperture": 0.05}, ()),
]


def _mixed_passes():
    from rs_amd.scenes import sky_env_map
    out = []
    for props, w, h, seed, env, _, _ in MIXED:
        p = rs_amd.ReservoirSplatting({**props, "maxSurfaceBounces": 4})
        p.setScene(scenes.box_room(seed=seed))
        if env:
            p.setEnvMap(sky_env_map(32, 16), intensity=0.5)
        out.append(p)
    return out


def _mixed_io(f):
    """Per-pass outputs of frame f (device tensors) and the caller's motion vectors."""
    ios = []
    for i, (_, w, h, _, _, _, outs) in enumerate(MIXED):
        d = {}
        if "rayCount" in outs:
            d["rayCount"] = torch.zeros((h, w), dtype=torch.int32, device="cuda")
        if "pathLength" in outs:
            d["pathLength"] = torch.full((h, w), 7, dtype=torch.int32, device="cuda")
        if "mvec" in outs:
            g = torch.Generator().manual_seed(100 * f + i)
            d["mvec"] = (torch.rand((h, w, 2), generator=g) * 4.0 - 2.0).cuda()
        ios.append(d)
    return ios


@pytest.mark.parametrize("form,count", [("fused", 1), ("fused", 3), ("streams", 2)])
def test_batch_mixed_contexts(form, count):
    """Every option branch the fused form reads from its tables, against each pass's own frames."""
    print(f"mixed batch {form} {count}", flush=True)
    batched, single = _mixed_passes(), _mixed_passes()
    batch = Batch(batched)
    _configure(batch, form, count, 16)
    sizes = [(w, h) for _, w, h, _, _, _, _ in MIXED]
    names = ("history_reservoirs", "history_recon", "temporal_reservoirs", "temporal_recon", "color", "mvec")
    for f in range(5):
        if f == 2:  # a frozen scene keeps frameCount and the history (rs_set_frozen)
            batched[0].frozen = True
            single[0].frozen = True
        if f == 4:
            batched[0].frozen = False
            single[0].frozen = False
        cams = [scenes.box_room_camera(f, w, h, focal_distance=3.5, **kw) for (_, w, h, _, _, kw, _) in MIXED]
        ios_b, ios_s = _mixed_io(f), _mixed_io(f)
        colors_b = [torch.zeros((h, w, 4), dtype=torch.float32, device="cuda") for w, h in sizes]
        batch.render(cams, sizes, colors=colors_b, ios=ios_b)
        for p, cam, (w, h), io in zip(single, cams, sizes, ios_s):
            p.execute(cam, w, h, **io)
        torch.cuda.synchronize()
        for i, (p, q, (w, h)) in enumerate(zip(batched, single, sizes)):
            for name in names:
                assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), (f, i, name)
            if MIXED[i][0].get("captureStages"):
                for name in ("initial_reservoirs", "initial_recon"):
                    assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), (f, i, name)
            assert np.array_equal(colors_b[i].cpu().numpy().ravel().view(np.uint8), q.read_buffer("color", w, h)), (f, i)
            for key in ("rayCount", "pathLength"):
                if key in ios_b[i]:
                    assert torch.equal(ios_b[i][key], ios_s[i][key]), (f, i, key)
            assert p.pixelStats == q.pixelStats, (f, i)
    batch.close()


@pytest.mark.parametrize("form,count", [("fused", 2), ("streams", 2)])
def test_batch_frames_on_alternating_streams(form, count):
    """Consecutive batch frames launched on different streams still see each other's history:
    each batch frame waits for the previous one."""
    batched, single = _passes(), _passes()
    batch = Batch(batched)
    _configure(batch, form, count, 16)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    sizes = [(w, h) for _, w, h, _ in SETUPS]
    for f in range(5):
        cams = [scenes.box_room_camera(f, w, h) for w, h in sizes]
        batch.render(cams, sizes, stream=streams[f % 2])
        for p, cam, (w, h) in zip(single, cams, sizes):
            p.execute(cam, w, h)
    torch.cuda.synchronize()
    for p, q, (w, h) in zip(batched, single, sizes):
        for name in ("history_reservoirs", "history_recon", "color"):
            assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), name
    batch.close()


@pytest.mark.parametrize("form,count", [("fused", 2), ("streams", 2)])
def test_batch_topology_changes(form, count):
    """A pass's options and a frame size change between batch frames: the batch records its
    launches again (the fused form builds a new topology) and stays equal to single frames."""
    batched, single = _passes(), _passes()
    batch = Batch(batched)
    _configure(batch, form, count, 16)
    sizes = [(w, h) for _, w, h, _ in SETUPS]
    for f in range(6):
        if f == 2:
            for p in (batched[1], single[1]):
                p.setProperties({"temporalReuse": "MultiScatter", "numTimePartitions": 2})
        if f == 4:
            sizes[2] = (56, 40)  # a new size resets that pass's history
        shutter = [0.0, 1.0 / 24.0, 0.0]
        cams = [scenes.box_room_camera(f, w, h, shutter=sh) for (w, h), sh in zip(sizes, shutter)]
        batch.render(cams, sizes)
        for p, cam, (w, h) in zip(single, cams, sizes):
            p.execute(cam, w, h)
        for i, (p, q, (w, h)) in enumerate(zip(batched, single, sizes)):
            for name in ("history_reservoirs", "history_recon", "color"):
                assert np.array_equal(p.read_buffer(name, w, h), q.read_buffer(name, w, h)), (f, i, name)
    batch.close()


def test_batch_autotune():
    batched = _passes()
    batch = Batch(batched)
    sizes = [(w, h) for _, w, h, _ in SETUPS]
    cams = [scenes.box_room_camera(0, w, h) for w, h in sizes]
    q, t, ms = batch.autotune(cams, sizes, frames=1)
    assert q in (0, 1, 2, 4, 8) and t in (8, 16, 32, 64) and ms > 0
    cfg = batch.config()
    assert cfg["queues"] == q and cfg["tile_width"] == t and 1 <= cfg["groups"] <= 8
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
