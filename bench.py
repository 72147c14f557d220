"""Benchmark of the per-frame dense SLAM path (track + fuse + raycast).

Workload: the 640x480 sequence of the seeded procedural room (fx = fy = 525,
c = (319.5, 239.5)), uint16 depth at 1 mm with sensor noise and 2 % dropout,
uint8 RGB and a 10-class label image fused as per-voxel labels; 5 mm voxels,
15 mm truncation, 2^20 buckets x 2, a 400k-block heap; ICP 10/5/3, FP32.
One step = one frame (System::GrabImage). All frames are resident in HBM
before the timed region; W warm-up frames run first, then K frames are timed
between two synchronisations. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOAD = dict(cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.005, trunc=0.015,
                buckets=1 << 20, blocks=400000, seed=0)
STAGES = ["pyramid", "track", "allocate", "visibility", "fuse", "prepass", "raycast"]
HBM_PEAK_GBS = 8000.0
# Bytes of one fused voxel: 8 B geometry + 12 B semantic record read and
# written, plus its depth (4), RGB (3), normal (16) and label (1) gathers.
FUSE_BYTES_PER_VOXEL = 2 * (8 + 12) + 4 + 3 + 16 + 1
FUSE_BYTES_PER_BLOCK = 16  # visible-list record


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(n, w):
    from semantic_icpslam_amd import binding as B
    syn = B.Synth(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], seed=w["seed"])
    frames = []
    for i in range(n):
        frames.append(syn.render(i))
    return syn, frames


def cpu_baseline(w, frames, seconds=20.0):
    """The oracle (scalar C++ restatement, OpenMP over pixels and blocks) on the
    same frames, on all host cores, for a bounded prefix of the sequence."""
    from oracle import binding as O
    from semantic_icpslam_amd import binding as B
    cores = O.set_threads(0)
    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    sysc = O.OracleSystem(cfg)
    t0 = time.perf_counter()
    times = []
    for i, (d, rgb, lab) in enumerate(frames):
        t = time.perf_counter()
        sysc.grab(d, rgb, lab)
        times.append(time.perf_counter() - t)
        if time.perf_counter() - t0 > seconds:
            break
    n = len(times)
    # frame 0 has no tracking; report the mean over the tracked frames
    ms = 1000.0 * (np.mean(times[1:]) if n > 1 else times[0])
    return {"value": round(float(ms), 3), "unit": "ms/frame", "cores": int(cores), "kind": "port",
            "sample": f"frames 0-{n - 1} of the same 640x480 sequence (ms/frame over frames "
                      f"1-{n - 1}), oracle with {cores} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent sequences would shard one per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from semantic_icpslam_amd import binding as B

    w = WORKLOAD
    dev = torch.device("cuda:0")
    n = args.warmup + args.steps
    log(f"[bench] rendering {n} frames of {w['cols']}x{w['rows']}")
    syn, frames = make_frames(n, w)
    depth = [torch.from_numpy(d).to(dev) for d, _, _ in frames]
    rgb = [torch.from_numpy(c).to(dev) for _, c, _ in frames]
    lab = [torch.from_numpy(l).to(dev) for _, _, l in frames]
    torch.cuda.synchronize()

    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    sysg = B.System(cfg)
    stream = torch.cuda.current_stream()
    log(f"[bench] warmup {args.warmup} frames")
    for i in range(args.warmup):
        sysg.grab(depth[i], rgb[i], lab[i], stream)
    torch.cuda.synchronize()
    map_h = sysg.map_handle()
    st0 = B.map_stats(map_h, stream)

    # Timed region: K frames, stage events recorded on the same stream.
    B.check(B.lib().sicp_system_set_timing(sysg.h, 1))
    log(f"[bench] timing {args.steps} frames")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, n):
        sysg.grab(depth[i], rgb[i], lab[i], stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tot = np.zeros(len(STAGES))
    cnt = np.zeros(len(STAGES), np.int32)
    B.check(B.lib().sicp_system_read_timing(sysg.h, tot.ctypes.data, cnt.ctypes.data, len(STAGES),
                                            B.stream_ptr(stream)))
    B.check(B.lib().sicp_system_set_timing(sysg.h, 0))
    st1 = B.map_stats(map_h, stream)
    res = sysg.result(stream)
    gt = syn.pose(n - 1)
    ms_per_frame = 1000.0 * elapsed / args.steps
    stage_ms = {s: round(float(tot[k] / max(cnt[k], 1)), 4) for k, s in enumerate(STAGES)}
    log("[bench] per-stage ms/frame: " + json.dumps(stage_ms))

    # Roofline of the fusion kernel: algorithmic bytes of the timed frames
    # over its summed event time.
    upd = st1["updated_total"] - st0["updated_total"]
    vis = st1["visible_total"] - st0["visible_total"]
    fus = max(st1["fusions"] - st0["fusions"], 1)
    fuse_bytes = upd * FUSE_BYTES_PER_VOXEL + vis * FUSE_BYTES_PER_BLOCK
    fuse_s = tot[STAGES.index("fuse")] / 1000.0
    achieved = fuse_bytes / max(fuse_s, 1e-12) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "fuse_kernel",
                "bytes_per_launch": int(fuse_bytes / fus),
                "updated_voxels_per_launch": int(upd / fus),
                "visible_blocks_per_launch": int(vis / fus)}

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(w, frames, args.cpu_seconds)

    out = {
        "metric": "ms/frame track+fuse+raycast 640x480 (plain FP32)",
        "value": round(ms_per_frame, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_frame, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural room, seed 0)",
        "config": {"workload": "640x480 / 5 mm voxels / T 15 mm / 2^20x2 buckets / "
                               "400k-block heap / ICP 10-5-3 / RGB + 10-class labels",
                   "frames": n, "global_batch": 1},
        "stage_ms_per_frame": stage_ms,
        "allocated_blocks": st1["allocated_blocks"],
        "last_frame_tracked": res["tracked"],
        "last_frame_translation_vs_ground_truth_m": round(
            float(np.linalg.norm(res["pose"][:3, 3] - gt[:3, 3])), 5),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
