"""Fused frames/sec of the per-frame KinectFusion path on one MI355X.

Workload (the single-GPU headline configuration): the plain FP32 method at full
size — seeded synthetic 640x480 uint16 depth frames (seed 2: floor + 3 walls +
8 spheres, smooth trajectory <= 2 cm / 1 deg per frame, noise N(0,(0.0012 z^2)^2)),
fx = fy = 525, cx = 319.5, cy = 239.5, fused into a 512^3 FP32 TSDF + FP32
weights over a 4 m cube, truncation 3 cm, ICP 3 levels x (10, 5, 4) iterations,
distance 10 cm, normal angle 20 deg (no early stop).

A step = one frame: preprocess -> raycast L0..L2 at the previous pose -> ICP ->
fuse at the new pose.  Frame 0 (the initial fusion) and `--warmup` frames run
untimed; then exactly `--steps` frames are timed between two stream
synchronisations, with every frame's raw depth already resident in HBM.

Prints one JSON line (value = frames/s).  The roofline entry is the fuse's voxel
sweep (fuse_list_kernel): algorithmic bytes per launch = touched voxels x 16 B
(read f, w; write f, w) + 640*480*4 B of depth, divided by that kernel's average
duration measured with HIP events recorded around it on the pipeline's stream
inside the timed region.  "batched" and "serving" report the multi-session forms
(see their docstrings); both keep every input resident before timing.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "kinectfusion-cuda_amd"))

METRIC = "fused frames/sec (640x480→512³) plain vs batched; p99 frame latency @64 sessions"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=9)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-sessions", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-sessions", type=int, default=64)
    ap.add_argument("--serve-frames", type=int, default=150, help="frames per served session")
    ap.add_argument("--serve-sat-frames", type=int, default=60, help="frames per session in the saturating load")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (the known-size copy that "
                         "scales HBM counters in counter-collection runs); not part of the JSON line")
    return ap.parse_args()


def batched(args, kfusion, stream):
    """Sessions of 640x480 depth into 512^3 snorm16 + uint8 volumes over 4 m cubes (session s:
    seed 100 + s), advanced together: one tick = one frame of every session, each stage one
    launch across all sessions, each tick one hipGraph launch.  Launch shapes come from the
    autotuner during the untimed ticks.  Returns the batched entry of the JSON line."""
    import ctypes as C

    import numpy as np

    from kfusion.synth import SyntheticSequence
    lib = kfusion.lib()
    S, n, L, mue, W, H = args.batch_sessions, 512, 4.0, 0.03, 640, 480
    ticks = 1 + args.batch_warmup + args.batch_steps
    seqs = [SyntheticSequence(seed=100 + s, n_frames=max(ticks, 30), L=L) for s in range(S)]
    frames = np.stack([q.frames(0, ticks) for q in seqs])  # (S, ticks, H, W)
    dev = kfusion.DeviceArray.from_host(frames)
    b = kfusion.Batch(S, icp=kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0))), use_graph=True,
                      autotune=True, stream=stream)
    ids = [b.add(kfusion.SessionConfig(W, H, 525.0, 525.0, 319.5, 239.5, 1e-3, n, L, mue,
                                       kfusion.KF_VOXEL_SNORM16_U8, seqs[s].pose(0))) for s in range(S)]
    fb = W * H * 2

    def tick(t):
        b.tick(ids, [dev.ptr.value + (s * ticks + t) * fb for s in range(S)])

    # initial fusion, then untimed ticks (the autotuner's samples come from these)
    tick(0)
    for t in range(1, 1 + args.batch_warmup):
        tick(t)
    b.sync()
    extra = 0
    while not b.shape()[1] and extra < 40:  # finish tuning on repeated frames of the last warm-up tick
        b.tick(ids, [dev.ptr.value + (s * ticks + args.batch_warmup) * fb for s in range(S)])
        extra += 1
    b.sync()
    b.tick_times()
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    lib.kf_event_create(C.byref(ev0))
    lib.kf_event_create(C.byref(ev1))
    lib.kf_event_record(ev0, stream)
    for t in range(1 + args.batch_warmup, ticks):
        tick(t)
    lib.kf_event_record(ev1, stream)
    b.sync()
    gpu_ms = C.c_float()
    lib.kf_event_elapsed_ms(ev0, ev1, C.byref(gpu_ms))
    tot, nt, mx = b.tick_times()
    failed = sum(1 for i in ids if b.status(i) != 0)
    errs = []
    for s in range(S):
        p = b.pose(ids[s])
        errs.append(float(np.abs(p[12:15] - seqs[s].pose(ticks - 1)[12:15]).max()))
    shape, tuned = b.shape()
    sb = kfusion.session_bytes(kfusion.SessionConfig(W, H, volume_n=n, volume_size=L, mue=mue))
    b.close()
    dev.free()
    fps = S * args.batch_steps / (gpu_ms.value / 1e3)
    return {"value": round(fps, 3), "unit": "frames/s", "sessions": S, "ticks_timed": args.batch_steps,
            "ms_per_tick": round(gpu_ms.value / args.batch_steps, 4), "max_tick_ms": round(mx, 4),
            "workload": f"{S} sessions x 640x480 -> 512^3 snorm16+u8 over 4 m cubes, seeds 100..{99 + S}, "
                        "ICP 10/5/4 (FP32 sums), one hipGraph per tick",
            "dtype": "snorm16/u8 volume, f32 compute", "launch_shape": shape, "autotuned": tuned,
            "hbm_per_session_gb": round(sb / 1e9, 3), "icp_failures": failed,
            "final_pose_err_vs_truth_m": {"max": round(max(errs), 5), "median": round(float(np.median(errs)), 5),
                                          "seeds_over_5cm": [100 + s for s in range(S) if errs[s] > 0.05]}}


def measured_traffic():
    """HBM bytes per launch of the fuse sweep on this workload, from the counter runs summarised
    in profiles/fuse_list_kernel_traffic_plain512.md (FETCH_SIZE and WRITE_SIZE scaled by a
    known-size copy); None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "fuse_list_kernel_traffic_plain512.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch fuse_list_kernel<F32> plain-640x480-512:"):
            return int(line.split(":")[1])
    return None


def cpu_baseline(seq, raws, n, L, mue, budget_s):
    """The oracle (CPU restatement of the same path, OpenMP over pixels/voxels) on the
    first frames of the same sequence, bounded to about budget_s seconds."""
    import numpy as np
    from oracle.kfref import OraclePipeline, lib as oracle_lib
    threads = oracle_lib().kfo_threads()
    op = OraclePipeline(640, 480, n, L, mue, 525.0, 525.0, 319.5, 239.5, 1e-3, (10, 5, 4), 0.1,
                        float(np.deg2rad(20.0)), init_pose=seq.pose(0))
    op.frame(raws[0])
    t0 = time.perf_counter()
    done = 0
    for f in range(1, raws.shape[0]):
        op.frame(raws[f])
        done += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    op.close()
    return {"value": round(done / el, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 pipeline, frames 1..{done} of the same seed-2 sequence after the "
                      f"initial fusion, 512^3 volume, {el:.1f} s on {threads} OpenMP threads"}


def serving(args, kfusion, shape):
    """The serving loop under its Poisson load: sessions s = 0..S-1 (scene seed 1000 + s) of
    640x480 depth, volumes drawn from {256, 384, 512, 640, 768}^3 (seed 99) as snorm16 + uint8
    over 4 m cubes, admitted against a 200 GB HBM budget as they open (uniformly over the first
    second); each sends --serve-frames frames as a Poisson process of mean 30 Hz with +-10 %
    uniform jitter on every inter-arrival (seed 7), replaying a 30-frame bank ping-pong.
    Returns the serving entry of the JSON line (aggregate frames/s, p50/p99 latency)."""
    import numpy as np

    from kfusion.synth import SyntheticSequence
    S, L, mue, W, H = args.serve_sessions, 4.0, 0.03, 640, 480
    rng = np.random.RandomState(99)
    sizes = rng.choice([256, 384, 512, 640, 768], S)
    opens = np.sort(rng.uniform(0.0, 1.0, S))
    seqs = [SyntheticSequence(seed=1000 + s, n_frames=30, L=L) for s in range(S)]
    banks = [q.frames(0, 30) for q in seqs]
    cfgs = [kfusion.SessionConfig(W, H, 525.0, 525.0, 319.5, 239.5, 1e-3, int(sizes[s]), L, mue,
                                  kfusion.KF_VOXEL_SNORM16_U8, seqs[s].pose(0)) for s in range(S)]
    budget = 200 * 10 ** 9
    srv = kfusion.Server(S, budget, icp=kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0))),
                         use_graph=True, shape=shape)
    tickets, st = srv.run_load(cfgs, banks, [args.serve_frames] * S, opens, rate_hz=30.0, jitter=0.1, seed=7)
    srv.close()
    # capacity: the same sessions under a load far above what one GPU serves (1 kHz per session,
    # all open at once); frames wait in the per-session queues, so every tick is full
    srv = kfusion.Server(S, budget, icp=kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0))),
                         use_graph=True, shape=shape)
    _, sat = srv.run_load(cfgs, banks, [args.serve_sat_frames] * S, np.zeros(S), rate_hz=1000.0, jitter=0.1, seed=7)
    srv.close()
    return {"value": round(st["frames_per_s"], 3), "unit": "frames/s", "sessions": S,
            "frames": int(st["frames"]), "wall_s": round(st["wall_s"], 3),
            "latency_ms": {"p50": round(st["latency_p50_ms"], 3), "p99": round(st["latency_p99_ms"], 3),
                           "max": round(st["latency_max_ms"], 3), "mean": round(st["latency_mean_ms"], 3)},
            "offered_frames_per_s": round(st["offered_frames_per_s"], 1), "ticks": int(st["ticks"]),
            "mean_frames_per_tick": round(st["mean_frames_per_tick"], 2),
            "admitted": int(st["admitted"]), "queued_admissions": int(st["queued_admissions"]),
            "peak_resident_gb": round(st["peak_resident_bytes"] / 1e9, 2), "hbm_budget_gb": budget / 1e9,
            "icp_failed_sessions": int(st["failed_frames"]),
            "workload": f"{S} sessions x 640x480, volumes {{256..768}}^3 snorm16+u8 (seed 99), Poisson 30 Hz "
                        f"+-10 % jitter (seed 7), {args.serve_frames} frames each, 200 GB budget",
            "volume_sizes": [int(x) for x in sizes],
            "saturated": {"frames_per_s": round(sat["frames_per_s"], 3), "frames": int(sat["frames"]),
                          "mean_frames_per_tick": round(sat["mean_frames_per_tick"], 2),
                          "load": f"all {S} sessions open at once, 1 kHz +-10 % per session, "
                                  f"{args.serve_sat_frames} frames each"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import ctypes as C

    import numpy as np

    import kfusion
    from kfusion import _capi
    from kfusion.synth import SyntheticSequence

    kfusion.init(0)
    lib = kfusion.lib()
    n, L, mue, W, H = 512, 4.0, 0.03, 640, 480
    total = 1 + args.warmup + args.steps
    seq = SyntheticSequence(seed=2, n_frames=max(total, 300), L=L)
    raws = seq.frames(0, total)
    dev = kfusion.DeviceArray.from_host(raws)
    stream = C.c_void_p()
    kfusion.check(lib.kf_stream_create(C.byref(stream)), "stream")
    cfg = _capi.kf_pipeline_config()
    cfg.frame_width, cfg.frame_height, cfg.depth_scale = W, H, 1e-3
    cfg.fx, cfg.fy, cfg.cx, cfg.cy = 525.0, 525.0, 319.5, 239.5
    cfg.volume_n, cfg.volume_size, cfg.mue, cfg.voxel_format = n, L, mue, kfusion.KF_VOXEL_F32
    for i, v in enumerate(seq.pose(0)):
        cfg.initial_pose[i] = float(v)
    cfg.icp = kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0)))
    h = C.c_void_p()
    kfusion.check(lib.kf_pipeline_create(C.byref(cfg), stream, C.byref(h)), "pipeline")
    frame_bytes = W * H * 2

    def run(first, count):
        for f in range(first, first + count):
            kfusion.check(lib.kf_pipeline_frame(h, dev.ptr.value + f * frame_bytes), "frame")

    st = C.c_int()
    run(0, 1 + args.warmup)
    kfusion.check(lib.kf_pipeline_sync(h, C.byref(st)), "sync")
    vol = C.c_void_p()
    lib.kf_pipeline_volume(h, C.byref(vol))
    touched0 = C.c_int64()
    lib.kf_volume_touched_total(vol, C.byref(touched0))
    kfusion.check(lib.kf_pipeline_set_timing(h, 1), "timing")
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    lib.kf_event_create(C.byref(ev0))
    lib.kf_event_create(C.byref(ev1))
    lib.kf_stream_synchronize(stream)
    t0 = time.perf_counter()
    lib.kf_event_record(ev0, stream)
    run(1 + args.warmup, args.steps)
    lib.kf_event_record(ev1, stream)
    kfusion.check(lib.kf_pipeline_sync(h, C.byref(st)), "sync")
    wall = time.perf_counter() - t0
    gpu_ms = C.c_float()
    lib.kf_event_elapsed_ms(ev0, ev1, C.byref(gpu_ms))
    if st.value != 0:
        raise RuntimeError(f"ICP failure during the timed frames: {st.value}")
    touched1 = C.c_int64()
    lib.kf_volume_touched_total(vol, C.byref(touched1))
    fuse_ms, launches = C.c_float(), C.c_int64()
    lib.kf_pipeline_fuse_kernel_ms(h, C.byref(fuse_ms), C.byref(launches))
    stages = np.zeros(4, np.float32)
    nfr = C.c_int64()
    lib.kf_pipeline_stage_times(h, stages.ctypes.data_as(C.c_void_p), C.byref(nfr))
    pose = np.zeros(16, np.float32)
    lib.kf_pipeline_pose(h, pose.ctypes.data_as(C.c_void_p))
    gt = seq.pose(total - 1)

    elapsed_s = gpu_ms.value / 1e3
    fps = args.steps / elapsed_s
    touched_per = (touched1.value - touched0.value) / args.steps
    alg_bytes = touched_per * 16.0 + W * H * 4.0
    achieved = alg_bytes / (fuse_ms.value * 1e-3) / 1e9
    result = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(gpu_ms.value / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain FP32 per-frame path, 640x480 uint16 depth -> 512^3 FP32 TSDF+weights over a "
                               "4 m cube, mue 3 cm, ICP 10/5/4 iterations, seed-2 synthetic sequence",
                   "frames_timed": args.steps, "volume_n": n, "cube_m": L, "mue_m": mue, "frame": [W, H],
                   "wall_s": round(wall, 4),
                   "stage_ms_per_frame": {k: round(float(v) / max(1, nfr.value), 4)
                                          for k, v in zip(kfusion.STAGE_NAMES, stages)},
                   "final_pose_err_vs_truth_m": round(float(np.abs(pose[12:15] - gt[12:15]).max()), 5)},
        "roofline": {"bound": "hbm", "kernel": "fuse_list_kernel<F32> (voxel sweep; the brick cull and depth tile-max are two short launches before it)", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(), "avg_launch_ms": round(fuse_ms.value, 5),
                     "touched_voxels_per_launch": round(touched_per), "algorithmic_bytes_per_launch": round(alg_bytes)},
    }
    lib.kf_pipeline_destroy(h)
    dev.free()
    shape = None
    if not args.no_batch:
        bt = batched(args, kfusion, stream)
        bt["vs_plain"] = round(bt["value"] / result["value"], 3)
        result["batched"] = bt
        shape = bt["launch_shape"]
    if not args.no_serve:
        result["serving"] = serving(args, kfusion, shape)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(seq, raws, n, L, mue, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    if args.copy_mb > 0:
        nb = args.copy_mb << 20
        a, b = kfusion.DeviceArray(nb, np.uint8), kfusion.DeviceArray(nb, np.uint8)
        kfusion.check(lib.kf_memcpy_d2d(b.ptr, a.ptr, nb, None), "calibration copy")
        kfusion.check(lib.kf_device_synchronize(), "calibration copy")
        a.free()
        b.free()


if __name__ == "__main__":
    main()
