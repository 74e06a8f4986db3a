"""Benchmark of the ReservoirSplatting per-frame pass on one MI355X.

Workload: 1920x1080, 1 spp, ScatterBackup temporal reuse with Clamped gather and Balance
MIS, max depth 6, spatial reuse (1 neighbour, radius 30), on the ~1M-triangle instanced
field (700 icospheres + 8,000 boxes + 64 emissive triangles, seed 2) with the camera
orbiting 0.5 degrees per frame.  A step is one frame.  Inputs (scene, BVH, history) are
resident in HBM before the timed region; motion vectors are produced on the GPU.

Prints one JSON line: fps as `value`, Mrays/s (closest-hit + visibility queries counted
with logTraceRay semantics), per-stage device times, the roofline entry of the dominant
kernel (`roofline`, HBM bytes; `roofline_valu`, useful VALU lane-ops against the vector
issue peak, from the committed lane-utilisation counters), the same frames lit also by a
sky environment map (`sky`), the relMSE of a reuse frame against path-traced accumulations
(`quality`), the adjoint/forward ratio (`adjoint`, `adjoint_1080p`), shortened runs of the
MultiScatter workload with moving geometry (`multiscatter`) and of the 16-scene batch
(`batch16`), and the CPU baseline (the scalar oracle on the box's host cores, bounded sample).

Other workloads (--workload): sky (the same frame with the field also lit by a 2048x1024
sky environment map), multi (MultiScatter with moving geometry, motion blur and depth of
field), batch (16 scenes in one hipGraph launch), accumulate (AccumulatePass) and import
(scene ingestion of a ~1M-triangle pbrt + PLY file, then frames of the imported scene).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "reservoir-splatting_amd"))

METRIC = "1080p frames/s & Mrays/s at matched relMSE; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
R_BYTES, C_BYTES = 32, 112  # packed PathReservoir / ReconnectionData
COPY_PROBE_BYTES = 1 << 30  # 1 GiB read + 1 GiB written by the calibration copy
SKY_W, SKY_H, SKY_INTENSITY = 2048, 1024, 1.0  # the sky workload's lat-long environment map
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_1080p_field.json")
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_1080p_field.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-frames", type=int, default=3, help="frames of the CPU-baseline sample")
    ap.add_argument("--cpu-scale", type=int, default=1, help="CPU sample renders the frame at 1/scale per side")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the sky-lit frames and the relMSE comparison")
    ap.add_argument("--workload", choices=["field", "sky", "batch", "multi", "accumulate", "import"],
                    default="field",
                    help="field: the 1080p ScatterBackup frame (the headline metric); sky: the same frame with a "
                         "2048x1024 sky environment map added; batch: 16 uneven scenes "
                         "per frame through the batched hipGraph path; multi: the 1080p MultiScatter frame with "
                         "motion blur and depth of field over ~2M triangles in 256 moving instances; "
                         "accumulate: the AccumulatePass kernels on 1080p RGBA32F frames; import: scene ingestion "
                         "of the field as a pbrt file with a binary PLY, then frames of the imported scene")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one device copy of COPY_PROBE_BYTES first (the calibration kernel of the HBM-traffic "
                         "counter runs)")
    return ap.parse_args()


def scene_bytes(scene):
    n = scene.num_triangles
    nv = scene.positions.shape[0]
    return (n - 1) * 64 + n * 48 + nv * 24 + n * 12 + n * 4 + len(scene.materials) * 48


def kernel_bytes_per_pixel(stage):
    """Compulsory HBM bytes per pixel of each stage's kernel at the packed record sizes."""
    rc = R_BYTES + C_BYTES
    return {
        # writes the candidate reservoir + reconnection record
        "initial": rc,
        # reads mvec + prev record at round(prevPixel); writes backup record + floating coords
        "collect": 8 + rc + rc + 8,
        # reads prev record; writes splat destination, counts
        "reproject": rc + 4 + 4,
        # curr + backup + one splat (k ~ 1) records, mvec, floating coords, cell count/offset,
        # sorted index, 4 bilinear prev reservoirs; writes the record
        "temporal": rc + rc + (4 + rc) + 8 + 8 + 8 + 4 * R_BYTES + rc,
        # central + one neighbour record; writes the record
        "spatial": 2 * rc + rc,
        "resolve": R_BYTES + 16,
    }[stage]


def cpu_baseline(args, scene, scenes_mod):
    """Oracle frames on the host cores: a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_harness as oh
    import rs_amd

    cores = oh.set_threads(0)
    lib = rs_amd.load_library()
    o = rs_amd.abi.Options()
    lib.rs_default_options(C.byref(o))
    for k, v in workload_options().items():
        setattr(o, k, v)
    W, H = args.width // args.cpu_scale, args.height // args.cpu_scale
    p = oh.OraclePass(o)
    p.setScene(scene)
    times = []
    for f in range(args.cpu_frames):
        cam = scenes_mod.orbit_camera(f, W, H)
        t0 = time.perf_counter()
        p.execute(cam, W, H)
        times.append(time.perf_counter() - t0)
    temporal = times[1:] if len(times) > 1 else times
    med = float(np.median(temporal))
    px_ratio = (args.width * args.height) / (W * H)
    return {
        "value": 1.0 / (med * px_ratio),
        "unit": "frames/s",
        "cores": int(cores),
        "kind": "port",
        "sample": f"{args.cpu_frames} frames at {W}x{H} (1/{args.cpu_scale**2} of the pixels) of the same scene and "
                  f"options; frames 1..{args.cpu_frames - 1} run temporal reuse; per-frame time scaled by the pixel "
                  f"ratio {px_ratio:.0f}; median {med * 1e3:.0f} ms per sampled frame",
    }


def adjoint_timing(rs, scene, W, H, frame_ms, iters=10):
    """Time the adjoint of the frozen last frame: the replay backward (gradients of an L2
    image loss with respect to every material's base colour and roughness) against the
    frame's forward time."""
    import torch
    rs.freeze()
    dev = torch.device("cuda")
    theta = torch.from_numpy(rs.materialParams()).to(dev)
    image = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    grad_image = torch.ones((H, W, 4), dtype=torch.float32, device=dev)
    grad = torch.empty(theta.shape, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    rs.replay_forward(theta, image, stream)
    rs.replay_backward(theta, grad_image, grad, stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(iters):
        rs.replay_forward(theta, image, stream)
    ev[1].record(stream)
    for _ in range(iters):
        rs.replay_backward(theta, grad_image, grad, stream)
    ev[2].record(stream)
    torch.cuda.synchronize()
    fwd = ev[0].elapsed_time(ev[1]) / iters
    bwd = ev[1].elapsed_time(ev[2]) / iters
    return {"materials": int(theta.shape[0]), "replay_forward_ms": fwd, "adjoint_ms": bwd, "frame_ms": frame_ms,
            "adjoint_forward_ratio": bwd / frame_ms}


def differentiable_workload(iters=10):
    """The differentiable form at its own size: the 512x512 box room with 64 materials
    (seed 6 parameters), ScatterBackup, 4 bounces; the forward frame time (median of 20
    frames after 4 warm-up frames), then the adjoint of the frozen last frame against an L2
    loss to a target rendered by the pass from the seed-5 parameters (8 accumulated frames)."""
    import numpy as np
    import torch

    import rs_amd
    from rs_amd import scenes

    W = H = 512
    props = {"temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "maxSurfaceBounces": 4}
    cam = scenes.box_room_camera(0, W, H)
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream()
    color = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    target = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    tgt = rs_amd.ReservoirSplatting(props)
    tgt.setScene(scenes.differentiable_room(5))
    acc = rs_amd.AccumulatePass({"precisionMode": "Double"})
    for _ in range(8):
        tgt.execute(cam, W, H, color=color, stream=stream)
        acc.execute(color, target, stream=stream)
    tgt.close()
    rs = rs_amd.ReservoirSplatting(props)
    scene = scenes.differentiable_room(6)
    rs.setScene(scene)
    for _ in range(4):
        rs.execute(cam, W, H, color=color, stream=stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    times = []
    for _ in range(20):
        ev[0].record(stream)
        rs.execute(cam, W, H, color=color, stream=stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        times.append(ev[0].elapsed_time(ev[1]))
    frame_ms = float(np.median(times))
    rs.freeze()
    theta = torch.from_numpy(rs.materialParams()).to(dev)
    image = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    grad = torch.empty(theta.shape, dtype=torch.float64, device=dev)
    rs.replay_forward(theta, image, stream)
    grad_image = (2.0 * (image - target)).contiguous()
    rs.replay_backward(theta, grad_image, grad, stream)
    torch.cuda.synchronize()
    t = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t[0].record(stream)
    for _ in range(iters):
        rs.replay_forward(theta, image, stream)
    t[1].record(stream)
    for _ in range(iters):
        rs.replay_backward(theta, grad_image, grad, stream)
    t[2].record(stream)
    torch.cuda.synchronize()
    fwd = t[0].elapsed_time(t[1]) / iters
    bwd = t[1].elapsed_time(t[2]) / iters
    loss = float(((image.double() - target.double()) ** 2).sum())
    rs.close()
    return {"workload": "512x512 box room, 64 materials (seed 6), target rendered from seed 5 (8 frames), "
                        "ScatterBackup, max depth 4", "materials": int(theta.shape[0]), "frame_ms": frame_ms,
            "replay_forward_ms": fwd, "adjoint_ms": bwd, "adjoint_forward_ratio": bwd / frame_ms, "loss": loss,
            "grad_norm": float(grad.norm())}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter runs (FETCH_SIZE + WRITE_SIZE,
    each scaled by what the counter reads for the calibration copy), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else float(k["hbm_bytes"])


def measured_valu_busy(kernel):
    """Fraction of SIMD cycles issuing vector ALU work over `kernel`'s duration, from the committed
    counter run (SQ_ACTIVE_INST_VALU x 4 / 1024 SIMDs / (GRBM_GUI_ACTIVE / 8 XCDs)), or None.  The
    tracing kernels are bound by VALU issue in the BVH walk, which the HBM fraction does not show."""
    try:
        with open(COUNTERS_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else float(k["valu_busy"])


VALU_LANES_PEAK = 256 * 4 * 16 * 2.4e9  # lane-ops/s: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz (wave64 op = 4 cycles)


def valu_roofline(kernel, kernel_ms):
    """VALU roofline of a tracing kernel from the committed counter run: useful lane-ops per
    launch = SQ_INSTS_VALU x 64 x lane utilisation (SQ_THREAD_CYCLES_VALU /
    (SQ_ACTIVE_INST_VALU x 64): the mean fraction of a VALU instruction's lanes that are
    active), divided by the kernel's mean device time of this run, against the peak issue rate
    of single-cycle-per-quad 32-bit VALU work.  None without a lane-utilisation counter run."""
    try:
        with open(COUNTERS_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or "lane_util" not in k:
        return None
    lane_ops = float(k["valu_insts"]) * 64.0 * float(k["lane_util"])
    achieved = lane_ops / (kernel_ms * 1e-3)
    return {"kernel": kernel, "bound": "valu", "achieved": achieved / 1e12, "peak": VALU_LANES_PEAK / 1e12,
            "unit": "Tlane-op/s", "frac": achieved / VALU_LANES_PEAK, "lane_util": float(k["lane_util"]),
            "valu_busy": float(k["valu_busy"]), "lane_ops_per_launch": lane_ops}


def quality(rs_props, scene, W, H, frames=16, ref_frames=1024, pt_checkpoints=(1, 2, 4, 8, 16, 32, 64, 128)):
    """Image quality at matched relMSE: with the camera at rest (orbit frame 0), ImageCompare's
    relMSE against a `ref_frames` path-traced accumulation of the pass's frame after `frames`
    frames of reuse, and of path-traced accumulations of 1, 2, 4, ... frames; the path-traced
    frame count whose relMSE matches the pass's single frame is read off those measurements
    (geometric interpolation between the bracketing counts, not extrapolated: relMSE of
    sparse images does not fall as 1 / frames), and with the two frame times gives the
    equal-quality speedup."""
    import math

    import torch

    import rs_amd
    from rs_amd import metrics, scenes

    cam = scenes.orbit_camera(0, W, H)
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()

    def run(props, n, checkpoints=(), ref=None):
        p = rs_amd.ReservoirSplatting(props)
        p.setScene(scene)
        acc = rs_amd.AccumulatePass({"precisionMode": "Double"})
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        rel = {}
        ms_total = 0.0
        for k in range(1, n + 1):
            ev[0].record(stream)
            p.execute(cam, W, H, color=color, stream=stream)
            ev[1].record(stream)
            acc.execute(color, out, stream)
            torch.cuda.synchronize()
            ms_total += ev[0].elapsed_time(ev[1])
            if k in checkpoints:
                rel[k] = metrics.rel_mse(out.double().cpu().numpy(), ref)
        img = out.double().cpu().numpy(), color.double().cpu().numpy()
        p.close()
        return img, rel, ms_total / n

    pt_props = dict(rs_props, enableTemporalResampling=False, enableSpatialResampling=False)
    (ref, _), _, pt_ms = run(pt_props, ref_frames)
    _, pt_rel, _ = run(pt_props, max(pt_checkpoints), pt_checkpoints, ref)
    (rs_acc, rs_last), _, rs_ms = run(rs_props, frames)
    target = metrics.rel_mse(rs_last, ref)
    equal = None
    ks = sorted(pt_rel)
    for lo, hi in zip(ks, ks[1:]):
        if pt_rel[lo] >= target >= pt_rel[hi]:
            t = (math.log(pt_rel[lo]) - math.log(target)) / (math.log(pt_rel[lo]) - math.log(pt_rel[hi]))
            equal = math.exp(math.log(lo) + t * (math.log(hi) - math.log(lo)))
            break
    return {"frames": frames, "reference_frames": ref_frames, "relmse_pass_frame": target,
            "relmse_pass_accumulated": metrics.rel_mse(rs_acc, ref),
            "relmse_path_traced": {str(k): v for k, v in pt_rel.items()},
            "path_traced_frames_for_equal_relmse": equal,
            "equal_quality_note": None if equal is not None else f"beyond {ks[-1]} path-traced frames",
            "pass_ms_per_frame": rs_ms, "path_traced_ms_per_frame": pt_ms,
            "equal_quality_speedup": equal * pt_ms / rs_ms if equal is not None and rs_ms > 0 else None,
            "camera": "orbit frame 0, at rest"}


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """The box's achievable HBM bandwidth: a device-to-device copy of `nbytes` timed with HIP
    events, read plus written bytes per second (the achievable counterpart of the 8 TB/s
    vendor peak the roofline uses)."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        dst.copy_(src)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    del src, dst
    return {"bytes_per_copy": 2 * nbytes, "ms": ms, "achieved_GBs": 2 * nbytes / (ms * 1e-3) / 1e9,
            "vendor_peak_GBs": HBM_PEAK_GBS}


STAGE_FRAMES = 10


def stage_pass(rs, run, first, n):
    """Per-stage device times (HIP events) and rays per frame over `n` frames run by
    run(frame), with stage timing on: the stages then run in order on the caller's stream."""
    import numpy as np

    import rs_amd

    rs.set_timing(True)
    stage_ms = {s: [] for s in rs_amd.abi.STAGES}
    rays = []
    for f in range(first, first + n):
        run(f)
        s = rs.stats()  # synchronizes on this frame's last event
        for k, name in enumerate(rs_amd.abi.STAGES):
            stage_ms[name].append(s.msStage[k])
        rays.append(s.raysClosestHit + s.raysVisibility)
    rs.set_timing(False)
    return {k: float(np.mean(v)) for k, v in stage_ms.items()}, float(np.mean(rays))


def sky_frames(rs, scene, W, H, steps):
    """The same frames lit also by the 2048x1024 sky environment map (the field workload has
    no environment, which leaves a third of the pixels black): fps of `steps` frames."""
    import numpy as np
    import torch

    from rs_amd import scenes

    rs.setEnvMap(scenes.sky_env_map(SKY_W, SKY_H), intensity=SKY_INTENSITY)
    rs.reset()
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    for f in range(3):
        rs.execute(scenes.orbit_camera(f, W, H), W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    rays = 0
    for f in range(3, 3 + steps):
        rs.execute(scenes.orbit_camera(f, W, H), W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    s = rs.stats()
    rays = int(s.raysClosestHit + s.raysVisibility)
    miss = float((rs.read_buffer("vbuffer", W, H).view(np.uint32).reshape(H, W, 3)[..., 0] == 0xFFFFFFFF).mean())
    zero = float((color[..., :3].abs().sum(-1) == 0).float().mean())
    rs.clearEnvMap()
    return {"fps": steps / dt, "ms_per_frame": 1e3 * dt / steps, "frames": steps, "rays_per_frame": rays,
            "primary_miss_fraction": miss, "zero_pixel_fraction": zero,
            "env_map": f"{SKY_W}x{SKY_H} sky, intensity {SKY_INTENSITY}"}


def workload_properties():
    """The workload as pass properties (the reference's property names)."""
    return {"samplesPerPixel": 1, "maxSurfaceBounces": 6, "enableTemporalResampling": True,
            "temporalReuse": "ScatterBackup", "gatherOption": "Clamped", "scatterBackupMISOption": "Balance",
            "enableSpatialResampling": True, "numSpatialIterations": 1, "neighborCount": 1}


def workload_options():
    return dict(samplesPerPixel=1, maxSurfaceBounces=6, maxDiffuseBounces=6, maxSpecularBounces=6,
                maxTransmissionBounces=6, enableTemporalResampling=1, temporalReuse=2, gatherMechanism=1,
                scatterBackupMIS=0, enableSpatialResampling=1, numSpatialIterations=1, neighborCount=1)


def batch_main(args):
    """16 scenes of uneven resolution and size rendered per call as one hipGraph launch
    (rs_batch_render, autotuned form, group or queue count and tile width), against rendering
    the same contexts one after another with rs_render_frame."""
    print(json.dumps(batch_run(args)), flush=True)


def batch_run(args):
    """The batch workload's measurements as a bench line (batch_main)."""
    import numpy as np
    import torch

    import rs_amd
    from rs_amd import scenes
    from rs_amd.batch import Batch

    t0 = time.time()
    work = scenes.batch_workload()
    print(f"# batch workload: {len(work)} scenes, {sum(w * h for _, w, h in work) / 1e6:.1f} Mpixel, "
          f"{sum(s.num_triangles for s, _, _ in work)} triangles, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    def make():
        ps = []
        for sc, w, h in work:
            p = rs_amd.ReservoirSplatting(workload_properties())
            p.setScene(sc)
            ps.append(p)
        return ps

    sizes = [(w, h) for _, w, h in work]
    colors = [torch.empty((h, w, 4), dtype=torch.float32, device="cuda") for w, h in sizes]
    stream = torch.cuda.current_stream()
    cams = lambda f: [scenes.orbit_camera(f, w, h) for w, h in sizes]  # noqa: E731
    # sequential frames
    seq = make()
    for f in range(args.warmup):
        for p, cam, (w, h), col in zip(seq, cams(f), sizes, colors):
            p.execute(cam, w, h, color=col, stream=stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for f in range(args.warmup, args.warmup + args.steps):
        for p, cam, (w, h), col in zip(seq, cams(f), sizes, colors):
            p.execute(cam, w, h, color=col, stream=stream)
    torch.cuda.synchronize()
    seq_ms = 1e3 * (time.perf_counter() - t) / args.steps
    del seq
    # batched frames
    bat = make()
    batch = Batch(bat)
    frame = 0
    batch.render(cams(frame), sizes, colors, stream)
    frame += 1
    q, tw, tune_ms = batch.autotune(cams(frame), sizes, colors, frames=2, stream=stream)
    tuned = batch.config()
    frame += 1
    for _ in range(args.warmup):
        batch.render(cams(frame), sizes, colors, stream)
        frame += 1
    torch.cuda.synchronize()
    t = time.perf_counter()
    host = 0.0
    for _ in range(args.steps):
        h0 = time.perf_counter()
        batch.render(cams(frame), sizes, colors, stream)
        host += time.perf_counter() - h0
        frame += 1
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / args.steps
    host_ms = 1e3 * host / args.steps
    # both forms at fixed settings: fused (one launch per stage for all scenes) and streams
    forms = {}
    for name, qn, groups, tile in (("fused_1g_16", 0, 1, 16), ("fused_4g_16", 0, 4, 16), ("fused_8g_16", 0, 8, 16),
                                   ("streams_8q_32", 8, 0, 32)):
        if qn == 0:
            batch.configure_fused(groups, tile)
        else:
            batch.configure(qn, tile)
        batch.render(cams(frame), sizes, colors, stream)
        frame += 1
        torch.cuda.synchronize()
        t = time.perf_counter()
        host_f = 0.0
        for _ in range(max(3, args.steps // 2)):
            h0 = time.perf_counter()
            batch.render(cams(frame), sizes, colors, stream)
            host_f += time.perf_counter() - h0
            frame += 1
        torch.cuda.synchronize()
        n = max(3, args.steps // 2)
        form_ms = 1e3 * (time.perf_counter() - t) / n
        # host cost of a call with the GPU idle (no back-pressure from queued frames)
        idle = []
        for _ in range(3):
            torch.cuda.synchronize()
            h0 = time.perf_counter()
            batch.render(cams(frame), sizes, colors, stream)
            idle.append(time.perf_counter() - h0)
            frame += 1
        torch.cuda.synchronize()
        forms[name] = {"ms_per_step": form_ms, "host_ms_per_step": 1e3 * host_f / n,
                       "host_ms_idle_gpu": 1e3 * min(idle)}
        print(f"# {name}: {forms[name]}", file=sys.stderr, flush=True)
    mpix = sum(w * h for w, h in sizes) / 1e6
    batch.close()
    return {
        "metric": "batched frames/s (16 uneven scenes per frame)", "value": 1e3 / ms, "unit": "batch frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "16 scenes, resolutions from {320x240,640x480,1280x720,1920x1080}, 1e4-2e6 triangles "
                               "log-uniform (seeds 1000-1015), ScatterBackup/Clamped, 1 spp, max depth 6",
                   "mpixel_per_frame": mpix, "triangles": sum(s.num_triangles for s, _, _ in work)},
        "mpixel_per_s": mpix * 1e3 / ms, "sequential_ms_per_step": seq_ms, "speedup_vs_sequential": seq_ms / ms,
        "autotuned": {**tuned, "ms": tune_ms},
        "host_ms_per_step": host_ms, "forms": forms,
    }


def multi_main(args):
    """1080p MultiScatter with motion blur and depth of field (8 time partitions, shutter
    1/24 s, simulated frame time 1/48 s, aperture 0.02 focused at 10) on ~2M triangles:
    256 rigidly moving icospheres among 0.7M static triangles.  A step is one frame:
    rs_update_instances (one 3x4 transform per instance uploaded; vertices transformed,
    emissive triangles updated and the BVH refitted on the GPU) and the pass."""
    print(json.dumps(multi_run(args)), flush=True)


def multi_run(args):
    """The MultiScatter workload's measurements as a bench line (multi_main)."""
    import numpy as np
    import torch

    import rs_amd
    from rs_amd import scenes

    W, H, K = args.width, args.height, 8
    t0 = time.time()
    scene, vel = scenes.moving_field(seed=3)
    frames = args.warmup + args.steps
    xforms = [scenes.instance_transforms(scene, f, vel) for f in range(frames + STAGE_FRAMES)]
    print(f"# scene: {scene.num_triangles} triangles in {len(vel)} moving of {scene.num_instances} instances, "
          f"built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    props = dict(workload_properties(), temporalReuse="MultiScatter", numTimePartitions=K,
                 artificialFrameTime=1.0 / 48.0)
    rs = rs_amd.ReservoirSplatting(props)
    rs.setScene(scene)
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    cam = lambda f: scenes.orbit_camera(f, W, H, aperture=0.02, focal_distance=10.0, shutter=1.0 / 24.0)  # noqa: E731
    for f in range(args.warmup):
        rs.updateInstances(xforms[f])
        rs.execute(cam(f), W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for f in range(args.warmup, frames):
        rs.updateInstances(xforms[f])
        rs.execute(cam(f), W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t_start) / args.steps
    update_ms = []

    def step(f):
        t = time.perf_counter()
        rs.updateInstances(xforms[f])
        torch.cuda.synchronize()
        update_ms.append(1e3 * (time.perf_counter() - t))
        rs.execute(cam(f), W, H, color=color, stream=stream)

    stage_avg, rays = stage_pass(rs, step, frames, STAGE_FRAMES)
    return {
        "metric": "1080p MultiScatter frames/s with moving geometry", "value": 1e3 / ms, "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{W}x{H} MultiScatter K={K}, shutter 1/24 s, frame time 1/48 s, aperture 0.02 "
                               "focused at 10, 1 spp, max depth 6, spatial 1 neighbour, 256 moving icospheres among "
                               "0.7M static triangles (seed 3), orbit",
                   "triangles": scene.num_triangles},
        "mrays_per_s": rays * 1e3 / ms / 1e6, "rays_per_frame": rays,
        "update_ms": float(np.mean(update_ms)), "pass_ms": float(sum(stage_avg.values())), "stage_ms": stage_avg,
    }


def import_main(args):
    """Scene ingestion of a large file: the ~1M-triangle instanced field (seed 2) written as a
    binary little-endian PLY (positions and normals per vertex, one face list) referenced from
    a pbrt-v4 file with a camera and a diffuse material; timed: rs_builder parse + flatten
    (import_scene), the upload with the GPU BVH build (setScene), and frames of the imported
    scene through its own camera."""
    import tempfile

    import numpy as np
    import torch

    import rs_amd
    from rs_amd import importers, scenes

    W, H = args.width, args.height
    field = scenes.instanced_field(seed=2)
    tmp = tempfile.mkdtemp(prefix="rs_import_")
    ply = os.path.join(tmp, "field.ply")
    nv, nt = field.positions.shape[0], field.num_triangles
    vert = np.empty(nv, dtype=[("p", "<f4", 3), ("n", "<f4", 3)])
    vert["p"], vert["n"] = field.positions, field.normals
    face = np.empty(nt, dtype=[("c", "u1"), ("i", "<i4", 3)])
    face["c"], face["i"] = 3, field.indices.astype(np.int32)
    with open(ply, "wb") as f:
        f.write(("ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
                 "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nelement face %d\n"
                 "property list uchar int vertex_indices\nend_header\n" % (nv, nt)).encode())
        f.write(vert.tobytes())
        f.write(face.tobytes())
    cam = scenes.orbit_camera(0, W, H)
    fov = 2 * np.degrees(np.arctan(0.5 * cam.frameHeight / cam.focalLength))
    with open(os.path.join(tmp, "field.pbrt"), "w") as f:
        f.write("LookAt %r %r %r  %r %r %r  0 1 0\n" % (*[float(x) for x in cam.position], *[float(x) for x in cam.target]))
        f.write('Camera "perspective" "float fov" %r\nFilm "rgb" "integer xresolution" %d "integer yresolution" %d\n'
                % (float(fov), W, H))
        f.write('WorldBegin\nMaterial "diffuse" "rgb reflectance" [0.5 0.5 0.5]\n'
                'LightSource "infinite" "rgb L" [1 1 1]\nShape "plymesh" "string filename" "field.ply"\n')
    mb = os.path.getsize(ply) / 1e6
    t0 = time.perf_counter()
    imp = importers.import_scene(os.path.join(tmp, "field.pbrt"))
    t_import = time.perf_counter() - t0
    rs = rs_amd.ReservoirSplatting(workload_properties())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    imp.apply(rs)
    torch.cuda.synchronize()
    t_upload = time.perf_counter() - t0
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    icam = imp.camera(W, H)
    for _ in range(args.warmup):
        rs.execute(icam, W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rs.execute(icam, W, H, color=color, stream=stream)
    torch.cuda.synchronize()
    frame_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    for name in os.listdir(tmp):
        os.remove(os.path.join(tmp, name))
    os.rmdir(tmp)
    print(json.dumps({
        "metric": "scene ingestion: triangles imported per second (pbrt + binary PLY)",
        "value": nt / t_import, "unit": "triangles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": frame_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"~1M-triangle instanced field (seed 2) as a {mb:.0f} MB binary PLY in a pbrt-v4 "
                               f"file; then {W}x{H} ScatterBackup frames of the imported scene",
                   "triangles": nt, "vertices": nv},
        "import_s": t_import, "ply_mb_per_s": mb / t_import, "upload_and_bvh_s": t_upload,
        "imported_fps": 1e3 / frame_ms, "roofline": None,
    }), flush=True)


def accumulate_main(args):
    """AccumulatePass on 1920x1080 RGBA32F frames already in HBM: each precision mode
    accumulates `steps` frames after `warmup`, timed with HIP events on the launch stream.
    Streaming kernels, so the roofline is HBM: Single moves 64 B per pixel (frame and sum
    read, sum and output written), SingleCompensated and Double 96 B."""
    import torch

    import rs_amd

    W, H = args.width, args.height
    P = W * H
    stream = torch.cuda.current_stream()
    frames = [torch.rand((H, W, 4), device="cuda") for _ in range(4)]
    out = torch.empty((H, W, 4), device="cuda")
    modes = {}
    for mode, bpp in (("Single", 64), ("SingleCompensated", 96), ("Double", 96)):
        acc = rs_amd.AccumulatePass({"precisionMode": mode})
        for i in range(args.warmup):
            acc.execute(frames[i % 4], out, stream)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for i in range(args.steps):
            acc.execute(frames[i % 4], out, stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / args.steps
        gbs = bpp * P / (ms * 1e-3) / 1e9
        modes[mode] = {"ms_per_frame": ms, "bytes_per_launch": bpp * P, "achieved_GBs": gbs,
                       "frac": gbs / HBM_PEAK_GBS}
    single = modes["Single"]
    print(json.dumps({
        "metric": "accumulated 1080p frames/s", "value": 1e3 / single["ms_per_frame"], "unit": "frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": single["ms_per_frame"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"AccumulatePass on {W}x{H} RGBA32F frames (Single; all modes in 'modes')"},
        "roofline": {"kernel": "k_accum_single", "bound": "hbm", "achieved": single["achieved_GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": single["frac"], "traffic": None},
        "modes": modes,
    }), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding the frame across GPUs is not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "batch":
        return batch_main(args)
    if args.workload == "multi":
        return multi_main(args)
    if args.workload == "accumulate":
        return accumulate_main(args)
    if args.workload == "import":
        return import_main(args)
    import numpy as np
    import torch

    import rs_amd
    from rs_amd import scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    W, H = args.width, args.height
    t0 = time.time()
    scene = scenes.instanced_field(seed=2)
    print(f"# scene: {scene.num_triangles} triangles, {len(scene.materials)} materials, built in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    rs = rs_amd.ReservoirSplatting(workload_properties())
    t0 = time.time()
    rs.setScene(scene)
    print(f"# scene upload + GPU LBVH build {time.time() - t0:.2f}s", file=sys.stderr, flush=True)
    sky = args.workload == "sky"
    if sky:
        t0 = time.time()
        rs.setEnvMap(scenes.sky_env_map(SKY_W, SKY_H), intensity=SKY_INTENSITY)
        torch.cuda.synchronize()
        print(f"# {SKY_W}x{SKY_H} sky upload + importance map {time.time() - t0:.2f}s", file=sys.stderr, flush=True)
    if args.copy_probe:
        src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    color = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    frame = 0
    for _ in range(args.warmup):
        rs.execute(scenes.orbit_camera(frame, W, H), W, H, color=color, stream=stream)
        frame += 1
    torch.cuda.synchronize()
    print(f"# warmup done ({args.warmup} frames)", file=sys.stderr, flush=True)
    # Timed frames run as a caller runs them: no stage events, so the stages that do not need
    # the initial candidates overlap them on the pass's side stream.
    t_start = time.perf_counter()
    for i in range(args.steps):
        rs.execute(scenes.orbit_camera(frame, W, H), W, H, color=color, stream=stream)
        frame += 1
        if (i + 1) % 20 == 0:
            torch.cuda.synchronize()
            print(f"# step {i + 1}/{args.steps}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    fps = args.steps / elapsed
    stage_avg, rays = stage_pass(rs, lambda f: rs.execute(scenes.orbit_camera(f, W, H), W, H, color=color,
                                                          stream=stream), frame, STAGE_FRAMES)
    frame += STAGE_FRAMES
    mrays = rays * fps / 1e6
    # dominant kernel: the stage with the largest mean device time.  vbuffer, initial,
    # collect, reproject and resolve are one kernel each; temporal and spatial run a
    # flag/scan/scatter compaction, a persistent task kernel and a combine, named here by
    # the task kernel; the sort is a scan and two small kernels.
    dom = max((k for k in stage_avg if k != "sort"), key=lambda k: stage_avg[k])
    stage_kernel = {"vbuffer": "k_vbuffer", "initial": "k_initial", "collect": "k_collect",
                    "reproject": "k_reproject", "temporal": "k_temporal_tasks", "spatial": "k_spatial_tasks",
                    "resolve": "k_resolve"}[dom]
    P = W * H
    alg_bytes = kernel_bytes_per_pixel(dom if dom != "vbuffer" else "resolve") * P + scene_bytes(scene)
    achieved = alg_bytes / (stage_avg[dom] * 1e-3) / 1e9
    workload = ("1920x1080 ScatterBackup (Clamped gather, Balance MIS), 1 spp, max depth 6, "
                "spatial 1 neighbour r=30, ~1M-triangle instanced field (seed 2), orbit 0.5 deg/frame")
    if sky:
        workload += f", lit also by a {SKY_W}x{SKY_H} sky environment map (intensity {SKY_INTENSITY})"
    result = {
        "metric": METRIC if not sky else "1080p frames/s under an environment map", "value": fps,
        "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "width": W, "height": H, "triangles": scene.num_triangles},
        "mrays_per_s": mrays, "rays_per_frame": rays,
        "stage_ms_note": f"device time per stage from {STAGE_FRAMES} further frames with stage events, which "
                         "run the stages one after another",
        "bvh": {"nodes4": int(rs.stats().bvhNodes), "levels": int(rs.stats().bvhLevels)},
        "stage_ms": stage_avg,
        "roofline": {"kernel": stage_kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None if sky else measured_traffic(stage_kernel),
                     "alg_bytes_per_launch": alg_bytes,
                     "valu_busy": None if sky else measured_valu_busy(stage_kernel)},
        "hbm_copy": copy_bandwidth(),
    }
    if sky:
        result["primary_miss_fraction"] = float(
            (rs.read_buffer("vbuffer", W, H).view(np.uint32).reshape(H, W, 3)[..., 0] == 0xFFFFFFFF).mean())
    result["adjoint_1080p"] = adjoint_timing(rs, scene, W, H, ms_per_step)
    if not sky:
        result["adjoint"] = differentiable_workload()
        print("# differentiable workload done", file=sys.stderr, flush=True)
        result["roofline_valu"] = valu_roofline(stage_kernel, stage_avg[dom])
        result["zero_pixel_fraction"] = float((color[..., :3].abs().sum(-1) == 0).float().mean())
        if not args.no_extras:
            result["sky"] = sky_frames(rs, scene, W, H, 20)
            print("# sky frames done", file=sys.stderr, flush=True)
            result["quality"] = quality(workload_properties(), scene, W, H)
            print("# quality done", file=sys.stderr, flush=True)
            # the other configurations' workloads, shortened: MultiScatter with moving geometry
            # and the 16-scene batch (full runs: --workload multi / --workload batch)
            short = argparse.Namespace(**{**vars(args), "steps": 20, "warmup": 3})
            m = multi_run(short)
            result["multiscatter"] = {k: m[k] for k in ("value", "unit", "ms_per_step", "steps", "mrays_per_s",
                                                        "update_ms", "stage_ms")}
            result["multiscatter"]["workload"] = m["config"]["workload"]
            print("# multiscatter done", file=sys.stderr, flush=True)
            short = argparse.Namespace(**{**vars(args), "steps": 6, "warmup": 2})
            b = batch_run(short)
            result["batch16"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "steps", "mpixel_per_s",
                                                   "sequential_ms_per_step", "speedup_vs_sequential", "autotuned",
                                                   "host_ms_per_step", "forms")}
            result["batch16"]["workload"] = b["config"]["workload"]
            print("# batch done", file=sys.stderr, flush=True)
    if not args.no_cpu and not sky:
        try:
            result["cpu_baseline"] = cpu_baseline(args, scene, scenes)
        except Exception as e:  # report, never hide
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
