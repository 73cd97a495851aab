"""Fused frames/sec of the per-frame KinectFusion path on one MI355X.

Workload (the single-GPU headline configuration): the plain FP32 method at full
size -- seeded synthetic 640x480 uint16 depth frames (seed 2: floor + 3 walls +
8 spheres, smooth trajectory <= 2 cm / 1 deg per frame, noise N(0,(0.0012 z^2)^2)),
fx = fy = 525, cx = 319.5, cy = 239.5, fused into a 512^3 FP32 TSDF + FP32
weights over a 4 m cube, truncation 3 cm, ICP 3 levels x (10, 5, 4) iterations,
distance 10 cm, normal angle 20 deg (no early stop).

A step = one frame: preprocess -> raycast L0..L2 at the previous pose -> ICP ->
fuse at the new pose.  Frame 0 (the initial fusion) and `--warmup` frames run
untimed; then exactly `--steps` frames are timed between two stream
synchronisations, with every frame's raw depth already resident in HBM.

Prints one JSON line (value = frames/s).  `roofline` is the fuse's voxel sweep
(fuse_list_kernel): algorithmic bytes per launch = touched voxels x 16 B (read
f, w; write f, w) + 640*480*4 B of depth, divided by that kernel's average
duration from HIP events recorded around it on the pipeline's stream inside the
timed region.  `roofline_raycast` is the raycast (first pass + finishing pass,
3 levels): |S| x 4 B + sum of W_L*H_L*32 B, where |S| is the number of distinct
voxels the reference march's samples read (counted by the oracle on the model
after the timed frames).  "batched" and "serving" report the multi-session
forms (see their docstrings); every input is resident before timing.
"precision" runs the precision form on its 640x480 -> 256^3 workload and compares it
with the FP64 oracle of the same algorithm.  The oracle runs only in the CPU leg
at the end (cpu_leg): the CPU baseline, the |S| counts and the FP64 comparison;
no GPU figure is measured while it runs.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "kinectfusion-cuda_amd"))

METRIC = "fused frames/sec (640x480→512³) plain vs batched; p99 frame latency @64 sessions"
HBM_PEAK_GBS = 8000.0
W, H = 640, 480
LEVEL_PIXELS = W * H + (W // 2) * (H // 2) + (W // 4) * (H // 4)


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
    ap.add_argument("--batch-stage-ticks", type=int, default=4, help="extra ticks timed kernel by kernel")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-sessions", type=int, default=64)
    ap.add_argument("--serve-frames", type=int, default=150, help="frames per served session")
    ap.add_argument("--serve-sat-frames", type=int, default=60, help="frames per session in the saturating load")
    ap.add_argument("--serve-target-seconds", type=float, default=3.0,
                    help="length of the served 512^3 load offered near capacity")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-footprint", action="store_true", help="skip the oracle's raycast footprint counts")
    ap.add_argument("--no-early-stop", action="store_true",
                    help="skip the plain run with the reference application's ICP early stop")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the precision form's run against the FP64 oracle")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (the known-size copy that "
                         "scales HBM counters in counter-collection runs); not part of the JSON line")
    return ap.parse_args()


def rot_err_deg(pa, pb):
    """Angle (deg) between the rotations of two column-major poses."""
    import numpy as np
    A = np.asarray(pa, np.float64).reshape(4, 4).T[:3, :3]
    B = np.asarray(pb, np.float64).reshape(4, 4).T[:3, :3]
    M = A.T @ B
    w = 0.5 * np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    return float(np.degrees(np.arctan2(np.linalg.norm(w), (np.trace(M) - 1.0) / 2.0)))


class Events:
    """A pair of HIP events on a stream (destroyed explicitly)."""

    def __init__(self, kfusion):
        import ctypes as C
        self.lib = kfusion.lib()
        self.a, self.b = C.c_void_p(), C.c_void_p()
        kfusion.check(self.lib.kf_event_create(C.byref(self.a)), "event")
        kfusion.check(self.lib.kf_event_create(C.byref(self.b)), "event")

    def ms(self):
        import ctypes as C
        out = C.c_float()
        self.lib.kf_event_elapsed_ms(self.a, self.b, C.byref(out))
        return out.value

    def close(self):
        for e in (self.a, self.b):
            if e.value:
                self.lib.kf_event_destroy(e)
        self.a.value = self.b.value = None


def batched(args, kfusion, stream):
    """Sessions of 640x480 depth into 512^3 snorm16 + uint8 volumes over 4 m cubes (session s:
    seed 100 + s) in the precision form's arithmetic (Kahan-compensated FP32 ICP sums, FP64
    solve), advanced together: one tick = one frame of every session, each stage one launch across
    all sessions, each tick one hipGraph launch.  Launch shapes come from the autotuner during the
    untimed ticks.  After the timed ticks, --batch-stage-ticks more ticks run kernel by kernel
    (direct launches with events between the kernels) for the per-kernel times and rooflines."""
    import ctypes as C

    import numpy as np

    from kfusion.synth import SyntheticSequence
    lib = kfusion.lib()
    S, n, L, mue = args.batch_sessions, 512, 4.0, 0.03
    ticks = 1 + args.batch_warmup + args.batch_steps + args.batch_stage_ticks
    seqs = [SyntheticSequence(seed=100 + s, n_frames=max(ticks, 30), L=L) for s in range(S)]
    frames = np.stack([q.frames(0, ticks) for q in seqs])  # (S, ticks, H, W)
    dev = kfusion.DeviceArray.from_host(frames)
    icp = kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0)), compensated=True)
    b = kfusion.Batch(S, icp=icp, use_graph=True, autotune=True, stream=stream)
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
    ev = Events(kfusion)
    lib.kf_event_record(ev.a, stream)
    t_end = 1 + args.batch_warmup + args.batch_steps
    for t in range(1 + args.batch_warmup, t_end):
        tick(t)
    lib.kf_event_record(ev.b, stream)
    b.sync()
    gpu_ms = ev.ms()
    ev.close()
    tot, nt, mx = b.tick_times()
    # per-kernel times on further frames (direct launches, same results as the graph)
    stage = None
    if args.batch_stage_ticks > 0:
        kfusion.check(lib.kf_batch_set_stage_timing(b.h, 1), "stage timing")
        touched0 = b.touched_total()
        for t in range(t_end, ticks):
            tick(t)
        ms = np.zeros(7, np.float64)
        k = C.c_int64()
        kfusion.check(lib.kf_batch_stage_times(b.h, ms.ctypes.data_as(C.c_void_p), C.byref(k)), "stage times")
        touched = (b.touched_total() - touched0) / max(1, k.value)
        kfusion.check(lib.kf_batch_set_stage_timing(b.h, 0), "stage timing")
        per = ms / max(1, k.value)
        stage = dict(zip(["preprocess", "raycast_first_pass", "raycast_finishing_pass", "icp", "fuse_brick_cull",
                          "fuse_sweep", "tick"], [round(float(v), 4) for v in per]))
        alg = touched * 6.0 + S * W * H * 4.0
        fuse_roof = {"bound": "hbm", "kernel": "fuse_batch_list_kernel (all sessions' voxel sweep)",
                     "achieved": round(alg / (per[5] * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / (per[5] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "avg_launch_ms": round(float(per[5]), 4), "touched_voxels_per_launch": round(touched),
                     "algorithmic_bytes_per_launch": round(alg),
                     "bytes_per_unit": "6 B per touched voxel (snorm16 f + uint8 w, read and written) + 4 B per "
                                       "depth pixel of every session"}
    errs, rerr = [], []
    last = t_end - 1 + (args.batch_stage_ticks if stage else 0)
    for s in range(S):
        p = b.pose(ids[s])
        errs.append(float(np.abs(p[12:15] - seqs[s].pose(last)[12:15]).max()))
        rerr.append(rot_err_deg(p, seqs[s].pose(last)))
    failed = sum(1 for i in ids if b.status(i) != 0)
    # the models after the last tick of a sample of sessions, at their poses: the CPU leg counts
    # the reference march's footprint on them for the raycast roofline
    ray_models = None
    if stage and not args.no_footprint and not args.no_cpu_baseline:
        sample = list(range(0, S, max(1, S // 4)))[:4]
        ray_models = {"models": [(b.volume_download(ids[s])[0], b.pose(ids[s])) for s in sample], "sample": sample,
                      "sessions": S, "ms": float(per[1] + per[2])}
    shape, tuned = b.shape()
    sb = kfusion.session_bytes(kfusion.SessionConfig(W, H, volume_n=n, volume_size=L, mue=mue))
    b.close()
    dev.free()
    fps = S * args.batch_steps / (gpu_ms / 1e3)
    out = {"value": round(fps, 3), "unit": "frames/s", "sessions": S, "ticks_timed": args.batch_steps,
           "ms_per_tick": round(gpu_ms / args.batch_steps, 4), "max_tick_ms": round(mx, 4),
           "workload": f"{S} sessions x 640x480 -> 512^3 snorm16+u8 over 4 m cubes, seeds 100..{99 + S}, "
                       "ICP 10/5/4 with Kahan-compensated FP32 sums and FP64 solve, one hipGraph per tick",
           "dtype": "snorm16/u8 volume, f32 compute", "launch_shape": shape, "autotuned": tuned,
           "hbm_per_session_gb": round(sb / 1e9, 3), "icp_failures": failed,
           "final_pose_err_vs_truth": {"trans_m_max": round(max(errs), 5),
                                       "trans_m_median": round(float(np.median(errs)), 5),
                                       "rot_deg_max": round(max(rerr), 4),
                                       "rot_deg_median": round(float(np.median(rerr)), 4),
                                       "seeds_over_5cm": [100 + s for s in range(S) if errs[s] > 0.05]}}
    if stage:
        out["stage_ms_per_tick"] = stage
        out["roofline_fuse"] = fuse_roof
    return out, ray_models


def recorded_traffic():
    """HBM bytes per launch of the fuse sweep on this workload from the counter runs summarised in
    profiles/fuse_list_kernel_traffic_plain512.md (FETCH_SIZE and WRITE_SIZE scaled by a known-size
    copy), with the touched-voxel count of that run; None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "fuse_list_kernel_traffic_plain512.md")
    if not os.path.exists(path):
        return None, None
    traffic = touched = None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch fuse_list_kernel<F32> plain-640x480-512:"):
            traffic = int(line.split(":")[1])
        if line.startswith("touched_voxels_per_launch fuse_list_kernel<F32> plain-640x480-512:"):
            touched = int(line.split(":")[1])
    return traffic, touched


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(seq, raws, n, L, mue, budget_s):
    """The oracle (CPU restatement of the same path, OpenMP over pixels/voxels) on the first
    frames of the same sequence, bounded to about budget_s seconds on all host threads, then a
    single-thread figure on a few further frames (about budget_s / 3)."""
    import numpy as np
    from oracle.kfref import OraclePipeline, lib as oracle_lib
    ol = oracle_lib()
    threads = ol.kfo_threads()

    def run(first, budget, max_frames):
        op = OraclePipeline(640, 480, n, L, mue, 525.0, 525.0, 319.5, 239.5, 1e-3, (10, 5, 4), 0.1,
                            float(np.deg2rad(20.0)), init_pose=seq.pose(0))
        for f in range(first):
            op.frame(raws[f])
        base = [op.stage_ms(k) for k in range(4)]
        t0 = time.perf_counter()
        done = 0
        for f in range(first, min(raws.shape[0], first + max_frames)):
            op.frame(raws[f])
            done += 1
            if time.perf_counter() - t0 >= budget:
                break
        el = time.perf_counter() - t0
        stages = [(op.stage_ms(k) - base[k]) / max(1, done) for k in range(4)]
        op.close()
        return done, el, stages

    done, el, stages = run(1, budget_s, raws.shape[0])
    ol.kfo_set_threads(1)
    try:
        d1, e1, st1 = run(1, budget_s / 3.0, 3)
    finally:
        ol.kfo_set_threads(threads)
    names = ["preprocess", "raycast", "icp", "fuse"]
    return {"value": round(done / el, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_threads": os.cpu_count(),
            "stage_ms_per_frame": {k: round(v, 2) for k, v in zip(names, stages)},
            "single_thread": {"value": round(d1 / e1, 4), "unit": "frames/s", "frames": d1,
                              "stage_ms_per_frame": {k: round(v, 2) for k, v in zip(names, st1)}},
            "sample": f"oracle FP32 pipeline, frames 1..{done} of the same seed-2 sequence after the "
                      f"initial fusion, 512^3 volume, {el:.1f} s on {threads} OpenMP threads; single thread: "
                      f"frames 1..{d1}, {e1:.1f} s"}


def serving(args, kfusion, shape, batched_fps):
    """The serving loop.  (1) mixed volumes: sessions s = 0..S-1 (scene seed 1000 + s) of 640x480
    depth, volumes drawn from {256, 384, 512, 640, 768}^3 (seed 99) as snorm16 + uint8 over 4 m
    cubes, admitted against a 200 GB HBM budget as they open (uniformly over the first second);
    each sends --serve-frames frames as a Poisson process of mean 30 Hz with +-10 % uniform jitter
    on every inter-arrival (seed 7), replaying a 30-frame bank ping-pong.  (2) its capacity: the
    same sessions all open at once at 1 kHz each, so every tick is full (latency is then queueing
    time and is not reported).  (3) the headline form: S sessions of 640x480 -> 512^3 snorm16
    (seed 100 + s) offered Poisson at 80 % of the batched capacity, reporting aggregate frames/s
    with p50/p99 latency over the whole run and over its steady second half."""
    import numpy as np

    from kfusion.synth import SyntheticSequence
    S, L, mue = args.serve_sessions, 4.0, 0.03
    rng = np.random.RandomState(99)
    sizes = rng.choice([256, 384, 512, 640, 768], S)
    opens = np.sort(rng.uniform(0.0, 1.0, S))
    seqs = [SyntheticSequence(seed=1000 + s, n_frames=30, L=L) for s in range(S)]
    banks = [q.frames(0, 30) for q in seqs]
    cfgs = [kfusion.SessionConfig(W, H, 525.0, 525.0, 319.5, 239.5, 1e-3, int(sizes[s]), L, mue,
                                  kfusion.KF_VOXEL_SNORM16_U8, seqs[s].pose(0)) for s in range(S)]
    budget = 200 * 10 ** 9
    icp = kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0)), compensated=True)
    srv = kfusion.Server(S, budget, icp=icp, use_graph=True, shape=shape)
    tickets, st = srv.run_load(cfgs, banks, [args.serve_frames] * S, opens, rate_hz=30.0, jitter=0.1, seed=7)
    srv.close()
    srv = kfusion.Server(S, budget, icp=icp, use_graph=True, shape=shape)
    _, sat = srv.run_load(cfgs, banks, [args.serve_sat_frames] * S, np.zeros(S), rate_hz=1000.0, jitter=0.1, seed=7)
    srv.close()
    del banks
    # (3) 512^3 sessions near capacity
    tseqs = [SyntheticSequence(seed=100 + s, n_frames=30, L=L) for s in range(S)]
    tbanks = [q.frames(0, 30) for q in tseqs]
    tcfgs = [kfusion.SessionConfig(W, H, 525.0, 525.0, 319.5, 239.5, 1e-3, 512, L, mue, kfusion.KF_VOXEL_SNORM16_U8,
                                   tseqs[s].pose(0)) for s in range(S)]
    rate = 0.8 * batched_fps / S if batched_fps else 30.0
    nfr = max(10, int(round(rate * args.serve_target_seconds)))
    srv = kfusion.Server(S, budget, icp=icp, use_graph=True, shape=shape)
    _, tgt = srv.run_load(tcfgs, tbanks, [nfr] * S, np.zeros(S), rate_hz=rate, jitter=0.1, seed=7)
    tlat = srv.latencies()
    srv.close()
    # steady state: the frames completed after the first half (the first ticks capture the
    # tick graphs and fuse every new session's first frame)
    steady = tlat[len(tlat) // 2:]
    lat = lambda d: {"p50": round(d["latency_p50_ms"], 3), "p99": round(d["latency_p99_ms"], 3),
                     "max": round(d["latency_max_ms"], 3), "mean": round(d["latency_mean_ms"], 3)}
    return {"value": round(st["frames_per_s"], 3), "unit": "frames/s", "sessions": S,
            "frames": int(st["frames"]), "wall_s": round(st["wall_s"], 3), "latency_ms": lat(st),
            "offered_frames_per_s": round(st["offered_frames_per_s"], 1), "ticks": int(st["ticks"]),
            "mean_frames_per_tick": round(st["mean_frames_per_tick"], 2),
            "admitted": int(st["admitted"]), "queued_admissions": int(st["queued_admissions"]),
            "peak_resident_gb": round(st["peak_resident_bytes"] / 1e9, 2), "hbm_budget_gb": budget / 1e9,
            "icp_failed_sessions": int(st["failed_frames"]),
            "workload": f"{S} sessions x 640x480, volumes {{256..768}}^3 snorm16+u8 (seed 99), Poisson 30 Hz "
                        f"+-10 % jitter (seed 7), {args.serve_frames} frames each, 200 GB budget, Kahan ICP sums",
            "volume_sizes": [int(x) for x in sizes],
            "saturated": {"frames_per_s": round(sat["frames_per_s"], 3), "frames": int(sat["frames"]),
                          "mean_frames_per_tick": round(sat["mean_frames_per_tick"], 2),
                          "load": f"all {S} sessions open at once, 1 kHz +-10 % per session, "
                                  f"{args.serve_sat_frames} frames each (latency is queueing time: not reported)"},
            "near_capacity_512": {"frames_per_s": round(tgt["frames_per_s"], 3), "frames": int(tgt["frames"]),
                                  "steady_offered_frames_per_s": round(S * rate, 1),
                                  "offered_frames_per_s": round(tgt["offered_frames_per_s"], 1),
                                  "latency_ms": lat(tgt),
                                  "steady_latency_ms": {"p50": round(float(np.percentile(steady, 50)), 3),
                                                        "p99": round(float(np.percentile(steady, 99)), 3),
                                                        "max": round(float(np.max(steady)), 3)},
                                  "mean_frames_per_tick": round(tgt["mean_frames_per_tick"], 2),
                                  "icp_failed_sessions": int(tgt["failed_frames"]),
                                  "load": f"{S} sessions x 640x480 -> 512^3 snorm16+u8 (seeds 100..{99 + S}), all "
                                          f"open at once, Poisson {rate:.1f} Hz +-10 % per session (80 % of the "
                                          f"batched capacity), {nfr} frames each; frames_per_s and offered are over "
                                          "the whole run (first arrival to last completion), steady_* over the "
                                          "second half of the completions"}}


def copy_peak(kfusion, stream):
    """Achievable HBM bandwidth: a 2 GiB device-to-device copy (read + write bytes / time)."""
    import numpy as np
    lib = kfusion.lib()
    nb = 2 << 30
    a, b = kfusion.DeviceArray(nb, np.uint8), kfusion.DeviceArray(nb, np.uint8)
    ev = Events(kfusion)
    best = None
    for _ in range(3):
        lib.kf_event_record(ev.a, stream)
        kfusion.check(lib.kf_memcpy_d2d(b.ptr, a.ptr, nb, stream), "copy")
        lib.kf_event_record(ev.b, stream)
        kfusion.check(lib.kf_stream_synchronize(stream), "copy")
        ms = ev.ms()
        best = ms if best is None else min(best, ms)
    ev.close()
    a.free()
    b.free()
    return round(2.0 * nb / (best * 1e-3) / 1e9, 1)


PREC_FRAMES, PREC_N, PREC_L = 30, 256, 3.0


def precision_run(kfusion):
    """The precision form on its own workload: seed-1 frames, 640x480 -> 256^3 snorm16 + uint8
    counts over a 3 m cube, Kahan-compensated ICP sums and FP64 solve, 30 frames.  Returns what the
    CPU leg compares with the FP64 oracle of the same algorithm: every frame's pose, the final
    volume and the predicted vertex maps."""
    import numpy as np

    from kfusion.synth import SyntheticSequence
    seq = SyntheticSequence(seed=1, n_frames=PREC_FRAMES, L=PREC_L)
    raws = seq.frames()
    th = float(np.deg2rad(20.0))
    gp = kfusion.Pipeline(W, H, seq.fx, seq.fy, seq.cx, seq.cy, 1e-3, PREC_N, PREC_L, 0.03,
                          voxel_format=kfusion.KF_VOXEL_SNORM16_U8, initial_pose=seq.pose(0),
                          icp=kfusion.IcpConfig((10, 5, 4), W, H, 0.1, th, compensated=True))
    dev = kfusion.DeviceArray.from_host(raws)
    poses, status = [], 0
    for f in range(PREC_FRAMES):
        gp.frame(dev.ptr.value + f * W * H * 2)
        status = status or gp.sync()
        poses.append(gp.pose())
    F, Wt = gp.volume_download()
    maps = [gp.map(4, lvl) for lvl in range(3)]
    gp.close()
    dev.free()
    return {"seq": seq, "raws": raws, "poses": poses, "F": F, "W": Wt, "maps": maps, "status": status}


def cpu_leg(result, args, seq, raws, n, L, mue, plain_model, batch_models, prec):
    """Everything that runs the oracle, after the GPU legs: the CPU baseline (cpu_baseline), the
    reference march's footprint |S| on the GPU's models for the raycast rooflines (the oracle as a
    counter of algorithmic bytes), and the precision form against the FP64 oracle of the same
    algorithm (the oracle as the checker)."""
    import numpy as np
    from oracle.kfref import Oracle, OraclePipeline

    result["cpu_baseline"] = cpu_baseline(seq, raws, n, L, mue, args.cpu_seconds)
    o = Oracle()
    kis = [o.intrinsics(525.0, 525.0, 319.5, 239.5, lvl)[1] for lvl in range(3)]
    if plain_model is not None:
        F, pose, ray_ms = plain_model
        distinct, samples = o.raycast_footprint(F, n, L, mue, pose, kis, W, H)
        rms = ray_ms * 1e-3
        rb = distinct * 4.0 + LEVEL_PIXELS * 32.0
        result["roofline_raycast"] = {
            "bound": "hbm", "kernel": "raycast_kernel + raycast_finish_kernel (3 levels)",
            "achieved": round(rb / rms / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(rb / rms / 1e9 / HBM_PEAK_GBS, 5), "avg_launch_ms": round(ray_ms, 5),
            "distinct_voxels": distinct, "algorithmic_bytes_per_launch": round(rb),
            "samples_per_frame": samples, "samples_per_s": round(samples / rms, 1),
            "bytes_per_unit": "4 B per distinct voxel the reference march reads (f32 f, trilinear and gradient "
                              "stencils) + 32 B per pixel (vertex + normal)",
            "sample": "|S| counted by the oracle on the model after the timed frames at the final pose"}
    if batch_models is not None and "batched" in result:
        distinct = samples = 0
        for F, pose in batch_models["models"]:
            d, sm = o.raycast_footprint(F, n, L, mue, pose, kis, W, H)
            distinct += d
            samples += sm
        S, sample = batch_models["sessions"], batch_models["sample"]
        scale = S / len(sample)
        alg = distinct * scale * 2.0 + S * LEVEL_PIXELS * 32.0
        rms = batch_models["ms"] * 1e-3
        result["batched"]["roofline_raycast"] = {
            "bound": "hbm", "kernel": "raycast_batch_kernel + raycast_batch_finish_kernel (3 levels, all sessions)",
            "achieved": round(alg / rms / 1e9, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / rms / 1e9 / HBM_PEAK_GBS, 5), "avg_launch_ms": round(float(rms * 1e3), 4),
            "distinct_voxels_per_tick": round(distinct * scale), "algorithmic_bytes_per_launch": round(alg),
            "samples_per_s": round(samples * scale / rms, 1),
            "sample": f"|S| and samples counted by the oracle on sessions {sample}, scaled by {scale:g}",
            "bytes_per_unit": "2 B per distinct voxel read (snorm16 f) + 32 B per pixel (vertex + normal)"}
    if prec is not None:
        ps = prec["seq"]
        op = OraclePipeline(W, H, PREC_N, PREC_L, 0.03, ps.fx, ps.fy, ps.cx, ps.cy, 1e-3, (10, 5, 4), 0.1,
                            float(np.deg2rad(20.0)), init_pose=ps.pose(0), dtype=np.float64, count_weights=True)
        rot, trans = [], []
        for f in range(PREC_FRAMES):
            op.frame(prec["raws"][f])
            a, b = np.asarray(prec["poses"][f], np.float64), np.asarray(op.pose(), np.float64)
            rot.append(rot_err_deg(a, b) * np.pi / 180.0)
            trans.append(float(np.abs(a[12:15] - b[12:15]).max()))
        F64, W64 = op.tsdf()
        both = (prec["W"] > 0) & (W64 > 0)
        dF = np.abs(prec["F"].astype(np.float64) - F64)[both]
        out = {"workload": f"seed-1 frames, 640x480 -> {PREC_N}^3 snorm16 + uint8 counts over "
                           f"{PREC_L:g} m, Kahan-compensated ICP sums, FP64 solve, {PREC_FRAMES} frames",
               "reference": "FP64 oracle of the same algorithm (count weights, unquantised F), run independently",
               "icp_status": int(prec["status"]),
               "pose_max_rot_rad": float(f"{max(rot):.3g}"), "pose_max_trans_m": float(f"{max(trans):.3g}"),
               "tsdf_rms_mue": float(f"{np.sqrt(np.mean(dF ** 2)):.3g}"),
               "tsdf_within_1e-3_mue": round(float((dF <= 1e-3).mean()), 5)}
        for lvl in range(3):
            vg, v64 = prec["maps"][lvl], op.map(1, lvl)
            ok = (vg[..., 3] == -1) & (v64[..., 3] == -1)
            dz = np.linalg.norm(vg[..., :3].astype(np.float64) - v64[..., :3], axis=-1)[ok]
            out[f"raycast_L{lvl}_within_1mm"] = round(float((dz <= 1e-3).mean()), 5)
            out[f"raycast_L{lvl}_mean_mm"] = round(1e3 * float(dz.mean()), 4)
        op.close()
        result["precision"] = out


def plain_config(kfusion, seq, n, L, mue, early_stop):
    """Pipeline configuration of the plain workload (640x480 -> n^3 FP32 over an L cube).  With
    early_stop, the reference application's ICP stop thresholds (main.cpp:58-59: 0.1 degree, 1 cm;
    icp.cpp:45-53 moves to the next level once an update changes the pose by less than both)."""
    import numpy as np

    from kfusion import _capi
    cfg = _capi.kf_pipeline_config()
    cfg.frame_width, cfg.frame_height, cfg.depth_scale = W, H, 1e-3
    cfg.fx, cfg.fy, cfg.cx, cfg.cy = 525.0, 525.0, 319.5, 239.5
    cfg.volume_n, cfg.volume_size, cfg.mue, cfg.voxel_format = n, L, mue, kfusion.KF_VOXEL_F32
    for i, v in enumerate(seq.pose(0)):
        cfg.initial_pose[i] = float(v)
    cfg.icp = kfusion.IcpConfig((10, 5, 4), W, H, 0.1, float(np.deg2rad(20.0)),
                                iteration_stop_thresh_angle=float(np.deg2rad(0.1)) if early_stop else 0.0,
                                iteration_stop_thresh_distance=0.01 if early_stop else 0.0)
    return cfg


def plain_early_stop(args, kfusion, stream, dev, seq, n, L, mue):
    """The plain workload again with the reference application's ICP early stop: the same frames,
    warm-up and timed count, in a pipeline of its own.  Levels end early once converged, and the
    iteration launches after the stop find the level done and return at once (the loop stays on
    the device), so ICP costs less.  Reported beside `value`, which runs every iteration."""
    import ctypes as C

    import numpy as np
    lib = kfusion.lib()
    cfg = plain_config(kfusion, seq, n, L, mue, True)
    h = C.c_void_p()
    kfusion.check(lib.kf_pipeline_create(C.byref(cfg), stream, C.byref(h)), "pipeline")
    fb = W * H * 2
    st = C.c_int()
    for f in range(1 + args.warmup):
        kfusion.check(lib.kf_pipeline_frame(h, dev.ptr.value + f * fb), "frame")
    kfusion.check(lib.kf_pipeline_sync(h, C.byref(st)), "sync")
    kfusion.check(lib.kf_pipeline_set_timing(h, 1), "timing")
    ev = Events(kfusion)
    lib.kf_event_record(ev.a, stream)
    for f in range(1 + args.warmup, 1 + args.warmup + args.steps):
        kfusion.check(lib.kf_pipeline_frame(h, dev.ptr.value + f * fb), "frame")
    lib.kf_event_record(ev.b, stream)
    kfusion.check(lib.kf_pipeline_sync(h, C.byref(st)), "sync")
    gpu_ms = ev.ms()
    ev.close()
    stages = np.zeros(4, np.float32)
    nfr = C.c_int64()
    lib.kf_pipeline_stage_times(h, stages.ctypes.data_as(C.c_void_p), C.byref(nfr))
    pose = np.zeros(16, np.float32)
    lib.kf_pipeline_pose(h, pose.ctypes.data_as(C.c_void_p))
    gt = seq.pose(args.warmup + args.steps)
    lib.kf_pipeline_destroy(h)
    if st.value != 0:
        raise RuntimeError(f"ICP failure during the early-stop frames: {st.value}")
    return {"value": round(args.steps / (gpu_ms / 1e3), 3), "unit": "frames/s", "frames_timed": args.steps,
            "ms_per_frame": round(gpu_ms / args.steps, 4),
            "icp": "up to 10/5/4 iterations, a level ends once an update moves the pose by < 0.1 degree and "
                   "< 1 cm (main.cpp:58-59, icp.cpp:45-53)",
            "stage_ms_per_frame": {k: round(float(v) / max(1, nfr.value), 4)
                                   for k, v in zip(kfusion.STAGE_NAMES, stages)},
            "final_pose_err_vs_truth": {"trans_m": round(float(np.abs(pose[12:15] - gt[12:15]).max()), 5),
                                        "rot_deg": round(rot_err_deg(pose, gt), 4)}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the path runs on one GPU: --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import ctypes as C

    import numpy as np

    import kfusion
    from kfusion.synth import SyntheticSequence

    kfusion.init(0)
    lib = kfusion.lib()
    n, L, mue = 512, 4.0, 0.03
    total = 1 + args.warmup + args.steps
    seq = SyntheticSequence(seed=2, n_frames=max(total, 300), L=L)
    raws = seq.frames(0, total)
    dev = kfusion.DeviceArray.from_host(raws)
    stream = C.c_void_p()
    kfusion.check(lib.kf_stream_create(C.byref(stream)), "stream")
    cfg = plain_config(kfusion, seq, n, L, mue, False)
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
    ev = Events(kfusion)
    lib.kf_stream_synchronize(stream)
    t0 = time.perf_counter()
    lib.kf_event_record(ev.a, stream)
    run(1 + args.warmup, args.steps)
    lib.kf_event_record(ev.b, stream)
    kfusion.check(lib.kf_pipeline_sync(h, C.byref(st)), "sync")
    wall = time.perf_counter() - t0
    gpu_ms = ev.ms()
    ev.close()
    if st.value != 0:
        raise RuntimeError(f"ICP failure during the timed frames: {st.value}")
    touched1 = C.c_int64()
    lib.kf_volume_touched_total(vol, C.byref(touched1))
    fuse_ms, launches = C.c_float(), C.c_int64()
    lib.kf_pipeline_fuse_kernel_ms(h, C.byref(fuse_ms), C.byref(launches))
    stages = np.zeros(4, np.float32)
    nfr = C.c_int64()
    lib.kf_pipeline_stage_times(h, stages.ctypes.data_as(C.c_void_p), C.byref(nfr))
    stage_ms = {k: float(v) / max(1, nfr.value) for k, v in zip(kfusion.STAGE_NAMES, stages)}
    pose = np.zeros(16, np.float32)
    lib.kf_pipeline_pose(h, pose.ctypes.data_as(C.c_void_p))
    gt = seq.pose(total - 1)

    elapsed_s = gpu_ms / 1e3
    fps = args.steps / elapsed_s
    touched_per = (touched1.value - touched0.value) / args.steps
    alg_bytes = touched_per * 16.0 + W * H * 4.0
    achieved = alg_bytes / (fuse_ms.value * 1e-3) / 1e9
    traffic, traffic_touched = recorded_traffic()
    result = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(gpu_ms / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain FP32 per-frame path, 640x480 uint16 depth -> 512^3 FP32 TSDF+weights over a "
                               "4 m cube, mue 3 cm, ICP 10/5/4 iterations, seed-2 synthetic sequence",
                   "frames_timed": args.steps, "volume_n": n, "cube_m": L, "mue_m": mue, "frame": [W, H],
                   "wall_s": round(wall, 4),
                   "stage_ms_per_frame": {k: round(v, 4) for k, v in stage_ms.items()},
                   "final_pose_err_vs_truth": {"trans_m": round(float(np.abs(pose[12:15] - gt[12:15]).max()), 5),
                                               "rot_deg": round(rot_err_deg(pose, gt), 4)}},
        "roofline": {"bound": "hbm", "kernel": "fuse_list_kernel<F32> (voxel sweep; the brick cull and depth "
                                               "tile-max are two short launches before it)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "recorded: rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch scaled by a known "
                                       "copy, profiles/fuse_list_kernel_traffic_plain512.md "
                                       f"({traffic_touched} touched voxels per launch in that run)",
                     "avg_launch_ms": round(fuse_ms.value, 5), "touched_voxels_per_launch": round(touched_per),
                     "algorithmic_bytes_per_launch": round(alg_bytes),
                     "bytes_per_unit": "16 B per touched voxel (f32 f + f32 w, read and written) + 4 B per depth "
                                       "pixel"},
    }
    plain_model = None
    if not args.no_footprint and not args.no_cpu_baseline:
        # the model the next frame's raycast would see, at the current pose (the CPU leg counts the
        # reference march's footprint on it)
        Fv = np.empty(n ** 3 * 2, np.float32)
        kfusion.check(lib.kf_volume_download(vol, Fv.ctypes.data_as(C.c_void_p)), "volume download")
        plain_model = (np.ascontiguousarray(Fv[0::2]), pose.copy(), stage_ms["raycast"])
        del Fv
    lib.kf_pipeline_destroy(h)
    if not args.no_early_stop:
        result["plain_reference_early_stop"] = plain_early_stop(args, kfusion, stream, dev, seq, n, L, mue)
    dev.free()
    result["hbm_copy_gbs"] = copy_peak(kfusion, stream)
    result["roofline"]["frac_of_copy"] = round(achieved / result["hbm_copy_gbs"], 4)
    shape = None
    batch_models = None
    if not args.no_batch:
        bt, batch_models = batched(args, kfusion, stream)
        bt["vs_plain"] = round(bt["value"] / result["value"], 3)
        result["batched"] = bt
        shape = bt["launch_shape"]
    prec = None
    if not args.no_precision and not args.no_cpu_baseline:
        prec = precision_run(kfusion)
    if not args.no_serve:
        sv = serving(args, kfusion, shape, result.get("batched", {}).get("value"))
        sv["near_capacity_512"]["steady_vs_plain"] = round(sv["near_capacity_512"]["steady_offered_frames_per_s"] /
                                                           result["value"], 3)
        result["serving"] = sv
    if not args.no_cpu_baseline:
        cpu_leg(result, args, seq, raws, n, L, mue, plain_model, batch_models, prec)
    print(json.dumps(result), flush=True)
    if args.copy_mb > 0:
        nb = args.copy_mb << 20
        a, b = kfusion.DeviceArray(nb, np.uint8), kfusion.DeviceArray(nb, np.uint8)
        kfusion.check(lib.kf_memcpy_d2d(b.ptr, a.ptr, nb, None), "calibration copy")
        kfusion.check(lib.kf_device_synchronize(), "calibration copy")
        a.free()
        b.free()
    kfusion.check(lib.kf_device_synchronize(), "final synchronisation")
    kfusion.check(lib.kf_stream_destroy(stream), "stream destroy")


if __name__ == "__main__":
    main()
