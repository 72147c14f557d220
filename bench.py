"""Benchmark of the per-frame dense SLAM path (track + fuse + raycast).

Workload: the 640x480 sequence of the seeded procedural room (fx = fy = 525,
c = (319.5, 239.5)), 300 frames of uint16 depth at 1 mm with sensor noise and
2 % dropout, uint8 RGB and a 10-class label image fused as per-voxel labels;
5 mm voxels, 15 mm truncation, 2^20 buckets x 2, a 400k-block heap; ICP
10/5/3, FP32. One step = one frame (System::GrabImage). Every pass runs the
whole sequence on a fresh system, so the map reaches its full size; the last
K frames (--steps) are timed between two synchronisations, after at least W
(--warmup) untimed frames. All frames are resident in HBM before the timed
region. Prints one JSON line.
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
SEQ_FRAMES = 300  # length of the 640x480 sequence
STAGES = ["pyramid", "track", "allocate", "visibility", "fuse", "prepass", "raycast",
          "unify", "frame", "icp_iteration", "track_finish"]
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (each input read once, each output written once).
# Fusion: per fused voxel the 8 B geometry and 12 B semantic records read and
# written, plus its depth (4), RGB (3), normal (16) and label (1) gathers;
# per visible block its list record.
FUSE_BYTES_PER_VOXEL = 2 * (8 + 12) + 4 + 3 + 16 + 1
FUSE_BYTES_PER_BLOCK = 16
# Raycast: per ray the zRange pair (8) and the vmap + nmap writes (32); per
# march step the voxel's sdf + weight (5); per interpolated read eight sdf
# values (32); per block lookup one hash entry (16).
RAY_BYTES_PER_RAY = 40
RAY_BYTES_PER_STEP = 5
RAY_BYTES_PER_INTERP = 32
RAY_BYTES_PER_LOOKUP = 16
# ICP: per pixel of a level, current vertex + normal and the gathered last
# vertex + normal (float4 each).
ICP_BYTES_PER_PIXEL = 64
# Allocation: one two-entry bucket (32 B) per table probe, the depth (4 B) of
# every pixel.
ALLOC_BYTES_PER_PROBE = 32
ALLOC_BYTES_PER_PIXEL = 4
# Visibility: the block record (16 B) of every allocated heap slot, the
# record of every visible block written to the list.
VIS_BYTES_PER_SLOT = 16
VIS_BYTES_PER_VISIBLE = 16
# Prepass: the visible list read (16 B a block), both zRange images written.
PREPASS_BYTES_PER_BLOCK = 16
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_vga_5mm_sequence.json")


def pyramid_bytes(cols, rows):
    """The frame pyramid's three launches per frame: (1) u16 depth and RGB in,
    raw + filtered depth and intensity out, the last frame's level-0 model
    maps in and levels 1-2 out; (2) filtered depth and intensity in, levels
    1-2 of both out; (3) per level depth and intensity in, vertex, normal and
    the two derivative images out."""
    p = [(cols >> lv) * (rows >> lv) for lv in range(3)]
    b1 = p[0] * (2 + 3 + 4 + 4 + 1 + 32) + 32 * (p[1] + p[2])
    b2 = p[0] * (4 + 1) + 5 * (p[1] + p[2])
    b3 = sum(q * (4 + 1 + 16 + 16 + 2 + 2) for q in p)
    return b1 + b2 + b3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(n, w):
    """Renders n frames of the seeded sequence, several at a time (the
    renderer is reentrant and releases the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    from semantic_icpslam_amd import binding as B
    syn = B.Synth(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], seed=w["seed"])
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        frames = list(ex.map(syn.render, range(n)))
    return syn, frames


def cpu_baseline(w, frames, first_timed, seconds=60.0):
    """The oracle (scalar C++ restatement, OpenMP over pixels and blocks) on the
    same frames, on all host cores: the whole sequence when it fits the time
    bound (its work counters over the GPU-timed frames then give the exact
    algorithmic bytes), else a prefix; also a one-thread time on a short
    sample."""
    from oracle import binding as O
    from semantic_icpslam_amd import binding as B
    cores = O.set_threads(0)
    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    sysc = O.OracleSystem(cfg)
    t0 = time.perf_counter()
    times = []
    work0 = None
    for i, (d, rgb, lab) in enumerate(frames):
        if i == first_timed:
            work0 = sysc.work()
        t = time.perf_counter()
        sysc.grab(d, rgb, lab)
        times.append(time.perf_counter() - t)
        if time.perf_counter() - t0 > seconds:
            break
    n = len(times)
    work1 = sysc.work()
    work = None
    if work0 is not None and n > first_timed:
        work = {k: work1[k] - work0[k] for k in work1}
        work["frames"] = n - first_timed
    del sysc
    # one thread, a few frames of the same sequence (fresh map)
    O.set_threads(1)
    sys1 = O.OracleSystem(cfg)
    t1 = []
    for i, (d, rgb, lab) in enumerate(frames[:4]):
        t = time.perf_counter()
        sys1.grab(d, rgb, lab)
        t1.append(time.perf_counter() - t)
    O.set_threads(cores)
    # frame 0 has no tracking; report the mean over the tracked frames
    ms = 1000.0 * (np.mean(times[1:]) if n > 1 else times[0])
    out = {"value": round(float(ms), 3), "unit": "ms/frame", "cores": int(cores), "kind": "port",
           "sample": f"frames 0-{n - 1} of the same 640x480 sequence (ms/frame over frames "
                     f"1-{n - 1}), oracle with {cores} OpenMP threads",
           "nproc": os.cpu_count(),
           "value_1_thread": round(float(1000.0 * np.mean(t1[1:])), 3),
           "sample_1_thread": "frames 1-3 of the same sequence, one thread"}
    return out, work


SMALL = dict(cols=320, rows=240, f=262.5, cx=159.5, cy=119.5, voxel=0.01, trunc=0.03,
             buckets=1 << 18, blocks=60000, seed=0)


def cpu_small():
    """The reference method on the CPU at the small size: 30 frames at
    320x240, 1 cm voxels (the oracle, all host cores and one thread), with
    the pose difference to the FP64 transcription."""
    from oracle import binding as O
    from semantic_icpslam_amd import binding as B
    w = SMALL
    _, frames = make_frames(30, w)
    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    cores = O.set_threads(0)
    out = {"workload": "320x240 / 1 cm voxels / T 3 cm / 2^18x2 buckets / 30 frames",
           "cores": int(cores)}
    poses = {}
    for label, threads, fp64 in (("all_cores", cores, False), ("one_thread", 1, False),
                                 ("fp64", cores, True)):
        O.set_threads(threads)
        sysc = O.OracleSystem(cfg, fp64=fp64)
        ts, ps = [], []
        for d, rgb, lab in frames:
            t = time.perf_counter()
            r = sysc.grab(d, rgb, lab)
            ts.append(1000.0 * (time.perf_counter() - t))
            ps.append(r["pose"])
        poses[label] = ps
        if not fp64:
            out[label] = {"ms_p50": round(float(np.percentile(ts[1:], 50)), 3),
                          "ms_p99": round(float(np.percentile(ts[1:], 99)), 3)}
    O.set_threads(cores)
    dt = [float(np.linalg.norm(a[:3, 3] - b[:3, 3])) for a, b in zip(poses["all_cores"], poses["fp64"])]
    out["pose_vs_fp64"] = {"max_translation_m": max(dt), "last_translation_m": dt[-1]}
    return out


def _rot_trans_err(a, b):
    """Rotation angle (rad) of b^-1 a and translation difference (m)."""
    dR = b[:3, :3].T @ a[:3, :3]
    return (float(np.arccos(np.clip((np.trace(dR) - 1) / 2, -1, 1))),
            float(np.linalg.norm(a[:3, 3] - b[:3, 3])))


def _download_image(im, shape, dtype):
    import ctypes
    import torch
    torch.cuda.synchronize()
    buf = np.zeros(shape, dtype)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    width = buf.strides[0]
    if hip.hipMemcpy2D(buf.ctypes.data, width, im.ptr, im.pitch, width, shape[0], 2) != 0:
        raise RuntimeError("hipMemcpy2D failed")
    return buf


def pose_precision(w, frames, syn, cfgs, sample_every=10):
    """Pose precision of the GPU system against FP64 (BASELINE's C3 report),
    for each configuration in `cfgs` (name -> system config), over the whole
    sequence (an untimed pass per configuration):
    - one step: at every `sample_every`-th frame, the GPU's tracked pose
      against the FP64 transcription of ComputeSE3 run on identical inputs —
      the frame's measured pyramid, the model pyramid the GPU tracked against
      (downloaded) and the GPU's previous pose;
    - drift: the GPU trajectory against the FP64 oracle system run on the
      same frames (each tracking against its own model), and each against the
      ground-truth trajectory, at the last frame."""
    import torch
    from oracle import binding as O
    from semantic_icpslam_amd import binding as B
    dev = torch.device("cuda:0")
    n = len(frames)
    cols, rows = w["cols"], w["rows"]
    first = next(iter(cfgs.values()))
    log("[bench] FP64 oracle system over the sequence")
    o64 = O.OracleSystem(first, fp64=True)
    ref = [o64.grab(d, rgb, lab)["pose"] for d, rgb, lab in frames]
    del o64
    out = {"what": "one_step: at every %d-th frame, the tracked pose against the FP64 "
                   "transcription of ComputeSE3 on the same measured pyramid, model pyramid and "
                   "previous pose (max / median over the samples); drift_vs_fp64: the trajectory "
                   "against the FP64 oracle system's over the same %d frames, at the last "
                   "frame; drift_vs_ground_truth: each at the last frame against the "
                   "synthetic trajectory" % (sample_every, n)}
    gt = syn.pose(n - 1)
    r, t = _rot_trans_err(ref[-1], gt)
    out["fp64_oracle_drift_vs_ground_truth"] = {"rotation_rad": r, "translation_m": t}
    for name, cfg in cfgs.items():
        g = B.System(cfg)
        steps, fp32_steps = [], []
        poses = []
        prev = None
        for i, (d, rgb, lab) in enumerate(frames):
            g.grab(torch.from_numpy(d).to(dev), torch.from_numpy(rgb).to(dev),
                   torch.from_numpy(lab).to(dev))
            res = g.result()
            poses.append(res["pose"])
            if i > 0 and i % sample_every == 0:
                model = []
                for lv in range(3):
                    shp = (rows >> lv, cols >> lv, 4)
                    model.append((_download_image(g.frame_image(16 + 0, lv), shp, np.float32),
                                  _download_image(g.frame_image(16 + 1, lv), shp, np.float32)))
                meas = O.measured_pyramid(d, w["f"], w["f"], w["cx"], w["cy"])
                p64, ok64 = O.track_frame(cfg, meas, model, prev, fp64=True)
                if ok64 and res["tracked"]:
                    steps.append(_rot_trans_err(res["pose"], p64))
                    p32, ok32 = O.track_frame(cfg, meas, model, prev, fp64=False)
                    if ok32:
                        fp32_steps.append(_rot_trans_err(p32, p64))
            prev = res["pose"]
        g.close()
        e = {}
        if steps:
            a = np.array(steps)
            e["one_step"] = {"samples": len(steps),
                             "max_rotation_rad": float(a[:, 0].max()),
                             "median_rotation_rad": float(np.median(a[:, 0])),
                             "max_translation_m": float(a[:, 1].max()),
                             "median_translation_m": float(np.median(a[:, 1]))}
        if fp32_steps:
            a = np.array(fp32_steps)
            e["one_step_fp32_oracle"] = {
                "what": "the FP32 transcription (the reference's arithmetic) on the same inputs",
                "max_rotation_rad": float(a[:, 0].max()),
                "max_translation_m": float(a[:, 1].max())}
        r, t = _rot_trans_err(poses[-1], ref[-1])
        e["drift_vs_fp64"] = {"frames": n, "rotation_rad": r, "translation_m": t}
        r, t = _rot_trans_err(poses[-1], gt)
        e["drift_vs_ground_truth"] = {"rotation_rad": r, "translation_m": t}
        out[name] = e
    return out


def precision_report(w, frames, syn, cfg32, cfg16, m=30):
    """The precision form against FP64 on the same frames: the first m frames
    fused at their ground-truth poses into an FP32-form map and a half-form
    map (GPU) and into the FP64 oracle map (CPU); TSDF RMS over the voxels
    all three hold (blocks matched by coordinate), and the raycast depth RMS
    at the last pose (one zRange for all three)."""
    import torch
    from oracle import binding as O
    from semantic_icpslam_amd import binding as B
    import copy
    dev = torch.device("cuda:0")
    K = B.intrinsics(w["f"], w["f"], w["cx"], w["cy"])
    mc = []
    for c in (cfg32.map, cfg16.map):  # a heap sized for m frames keeps the downloads small
        c = copy.copy(c)
        c.num_blocks = 40000
        mc.append(c)
    maps = [B.Map(mc[0]), B.Map(mc[1])]
    o64 = O.OracleMap(mc[0], fp64=True)
    for i, (d, rgb, lab) in enumerate(frames[:m]):
        raw, filt = O.filter_depth(d, w.get("scale", 1000.0), 3.0)
        nmap = O.compute_nmap(O.compute_vmap(filt, w["f"], w["f"], w["cx"], w["cy"], 3.0))
        T = syn.pose(i)
        pose = B.pose_f(T)
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (raw, rgb, nmap, lab)]
        for g in maps:
            B.check(B.lib().sicp_fuse_map_color(g.h, *[B.image(x) for x in t], 1, B.C.byref(pose),
                                                None, K, None, B.stream_ptr()))
        o64.fuse_fp64(raw, rgb, nmap, lab, 1, T, [w["f"], w["f"], w["cx"], w["cy"]])
    torch.cuda.synchronize()

    def voxels(exported, n):
        b = exported["blocks"][:n]
        return {(int(x), int(y), int(z)): int(sl) for x, y, z, sl in
                zip(b["x"], b["y"], b["z"], b["slot"])}, exported["geom"]

    d64 = o64.export()
    k64, g64 = voxels(d64, o64.stats()["allocated_blocks"])
    out = {"frames": m, "what": "fused at ground-truth poses; RMS over voxels held by all "
                                 "three maps with weight > 0"}
    dl = []
    for name, g in (("fp32", maps[0]), ("half", maps[1])):
        na = g.stats()["allocated_blocks"]
        e = g.download(n_blocks=na)
        kk, gg = voxels(e, na)
        common = [(kk[key], k64[key]) for key in kk.keys() & k64.keys()]
        a = np.concatenate([gg["sdf"][s * 512:(s + 1) * 512] for s, _ in common])
        wa = np.concatenate([gg["weight"][s * 512:(s + 1) * 512] for s, _ in common])
        b = np.concatenate([g64["sdf"][s * 512:(s + 1) * 512] for _, s in common])
        wb = np.concatenate([g64["weight"][s * 512:(s + 1) * 512] for _, s in common])
        ok = (wa > 0) & (wb > 0) & ~np.isnan(a) & ~np.isnan(b)
        diff = a[ok].astype(np.float64) - b[ok]
        rms = float(np.sqrt(np.mean(diff ** 2)))
        out[f"tsdf_rms_{name}"] = rms
        out[f"tsdf_rms_{name}_mm"] = rms * w["trunc"] * 1000.0
        out[f"tsdf_max_{name}"] = float(np.abs(diff).max())
        out["voxels_compared"] = int(ok.sum())
    # depth at the last pose, one zRange (from the FP32-form map) for all
    T = syn.pose(m - 1)
    pose = B.pose_f(T)
    zc, zr = (w["cols"] + 7) // 8, (w["rows"] + 7) // 8
    zmin = torch.zeros((zr, zc), dtype=torch.float32, device=dev)
    zmax = torch.zeros_like(zmin)
    ne = B.C.c_int(0)
    B.check(B.lib().sicp_create_rendering_blocks(maps[0].h, B.image(zmin), B.image(zmax), 3.0,
                                                 0.1, B.C.byref(pose), None, K, B.C.byref(ne),
                                                 B.stream_ptr()))
    invf = float(np.float32(1.0 / w["f"]))
    depths = []
    for g in maps:
        v = torch.empty((w["rows"], w["cols"], 4), dtype=torch.float32, device=dev)
        n = torch.empty_like(v)
        B.check(B.lib().sicp_raycast_inv(g.h, B.image(v), B.image(n), B.image(zmin),
                                         B.image(zmax), B.C.byref(pose), None, invf, invf,
                                         w["cx"], w["cy"], B.stream_ptr()))
        torch.cuda.synchronize()
        depths.append(v[..., 2].cpu().numpy().astype(np.float64))
    v64, _ = o64.raycast_fp64(w["cols"], w["rows"], zmin.cpu().numpy(), zmax.cpu().numpy(), T,
                              invf, invf, w["cx"], w["cy"])
    for name, dpt in zip(("fp32", "half"), depths):
        ok = ~np.isnan(dpt) & ~np.isnan(v64[..., 2])
        err = np.abs(dpt[ok] - v64[..., 2][ok]) * 1e3
        out[f"depth_rms_{name}_mm"] = float(np.sqrt(np.mean(err ** 2)))
        out[f"depth_p99_{name}_mm"] = float(np.percentile(err, 99))
        out[f"depth_max_{name}_mm"] = float(err.max())
        out[f"depth_pixels_over_1mm_{name}"] = int((err > 1.0).sum())
        out[f"depth_pixels_{name}"] = int(ok.sum())
        # hit / no-hit disagreements with the FP64 render
        out[f"depth_hit_mismatch_{name}"] = int((np.isnan(dpt) != np.isnan(v64[..., 2])).sum())
    for g in maps:
        g.close()
    return out


def roofline_entry(kernel, nbytes, seconds, launches, traffic_kib=None):
    achieved = nbytes / max(seconds, 1e-12) / 1e9
    e = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
         "bytes_per_launch": int(nbytes / max(launches, 1)),
         "avg_launch_us": round(1e6 * seconds / max(launches, 1), 2)}
    if traffic_kib is not None:
        e["traffic"] = int(traffic_kib * 1024)
    return e


def load_traffic():
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes, scaled by the same counters on a plain copy."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return {k: v["fetch_kib"] * t["copy_scale_fetch"] + v["write_kib"] * t["copy_scale_write"]
                for k, v in t["kernels"].items()}
    except (OSError, KeyError, ValueError):
        return {}


# Eight independent sequences of uneven resolution and voxel size.
BATCHED = (
    [dict(cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.005, seed=k) for k in (0, 1, 2)] +
    [dict(cols=320, rows=240, f=262.5, cx=159.5, cy=119.5, voxel=0.008, seed=k) for k in (3, 4, 5)] +
    [dict(cols=1280, rows=720, f=1050.0, cx=639.5, cy=359.5, voxel=0.004, seed=k) for k in (6, 7)])


def run_batched(args):
    """Eight independent sequences, each with its own system; one round
    advances every sequence by a frame. Passes over the same frames on fresh
    systems: the batch, a round as one hipGraph launch on one stream, in its
    two forms (every tracking iteration of all eight sequences in one launch;
    or each sequence's frame as an independent chain of the round), the
    faster of the two being the value; for comparison one stream per sequence
    with one hipGraph per frame, and all on one stream; and the value's batch
    form synchronised after every round (round latency)."""
    import torch
    from semantic_icpslam_amd import binding as B
    nfr = args.warmup + args.steps
    log(f"[bench] rendering {len(BATCHED)} sequences x {nfr} frames")
    seqs = []
    for w in BATCHED:
        w = dict(w, trunc=3 * w["voxel"], buckets=1 << 20, blocks=300000)
        _, frames = make_frames(nfr, w)
        dev = torch.device("cuda:0")
        seqs.append(dict(w=w, frames=[tuple(torch.from_numpy(a).to(dev) for a in f)
                                      for f in frames]))
    torch.cuda.synchronize()

    def run(mode, sync_rounds=False, tuning=None, autotune=False):
        """Fresh systems (raycast launch shapes from `tuning`); warm-up rounds,
        then the timed rounds. With `autotune`, the warm-up ends with every
        system's autotuner and the shapes it chose are returned instead."""
        systems, streams = [], []
        shared = torch.cuda.Stream()
        for k, q in enumerate(seqs):
            w = q["w"]
            cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"],
                                  w["voxel"], w["trunc"], w["buckets"], w["blocks"])
            cfg.icp_spread = 0  # eight sequences fill the GPU without spreading small levels
            systems.append(B.System(cfg))
            if tuning:
                systems[-1].set_tuning(tuning[k]["ray_first_pass_samples"],
                                       tuning[k]["ray_second_pass_workgroups"])
            streams.append(torch.cuda.Stream() if mode == "streams" else shared)
        batch = B.Batch(systems) if mode in ("batch", "chains") else None
        if mode == "chains":
            batch.set_shared_tracking(False)

        def round_(i):
            if batch:
                batch.grab_staged([q["frames"][i] for q in seqs], True, shared)
                return
            for q, g, st in zip(seqs, systems, streams):
                d, c, l = q["frames"][i]
                g.grab_staged(d, c, l, True, st)

        tuned = [None] * len(systems)
        for i in range(args.warmup):
            round_(i)
            if autotune and i > 0:  # sums over the warm-up rounds after the first
                for k, g in enumerate(systems):
                    try:
                        tuned[k] = g.autotune(repeats=3, stream=shared)
                    except B.SicpError as e:  # the frame was lost: skip this round
                        tuned[k] = tuned[k] or dict(g.tuning(), note=str(e))
        torch.cuda.synchronize()
        if autotune:
            if batch:
                batch.close()
            for g in systems:
                g.close()
            return tuned
        lat = []
        t0 = time.perf_counter()
        for i in range(args.warmup, nfr):
            t = time.perf_counter()
            round_(i)
            if sync_rounds:
                torch.cuda.synchronize()
                lat.append(1000.0 * (time.perf_counter() - t))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        tracked = [g.result(st)["tracked"] for g, st in zip(systems, streams)]
        if batch:
            batch.close()
        for g in systems:
            g.close()
        return elapsed, lat, tracked

    log("[bench] batched: autotuning each sequence's raycast launch shape after the warm-up")
    tuning = run("batch", autotune=True)
    log("[bench] batched: one launch per iteration for all sequences, a round per graph")
    shr_s, _, tracked = run("batch", tuning=tuning)
    log("[bench] batched: each sequence an independent chain of the round, a round per graph")
    chn_s, _, _ = run("chains", tuning=tuning)
    log("[bench] batched: one stream per sequence, one graph per frame")
    str_s, _, _ = run("streams", tuning=tuning)
    log("[bench] batched: one stream, one graph per frame")
    ser_s, _, _ = run("serial", tuning=tuning)
    best = "chains" if chn_s < shr_s else "batch"
    thr_s = min(chn_s, shr_s)
    log("[bench] batched: the batch synchronised per round")
    _, lat, _ = run(best, True, tuning=tuning)
    frames_done = args.steps * len(seqs)
    out = {
        "metric": "frames/s, 8 independent sequences (3x640x480 @5 mm, 3x320x240 @8 mm, "
                  "2x1280x720 @4 mm), a round of eight frames as one hipGraph launch on one "
                  "stream",
        "value": round(frames_done / thr_s, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * thr_s / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural rooms, seeds 0-7)",
        "config": {"workload": "batched8", "sequences": len(seqs), "frames_per_sequence": nfr,
                   "icp_spread": 0},
        "value_form": ("each sequence's frame an independent chain of the round's graph"
                       if best == "chains" else
                       "each Gauss-Newton iteration of all eight sequences in one launch"),
        "shared_iterations_frames_per_s": round(frames_done / shr_s, 1),
        "independent_chains_frames_per_s": round(frames_done / chn_s, 1),
        "streams_frames_per_s": round(frames_done / str_s, 1),
        "streams_what": "one stream per sequence, one hipGraph per frame",
        "one_stream_frames_per_s": round(frames_done / ser_s, 1),
        "round_ms_p50": round(float(np.percentile(lat, 50)), 4),
        "round_ms_p99": round(float(np.percentile(lat, 99)), 4),
        "round_what": "one round = every sequence advances one frame (step); p50/p99 from a "
                      "pass of the value's batch form synchronised after every round",
        "last_frame_tracked": tracked,
        "autotune": [dict(sequence=f"{q['w']['cols']}x{q['w']['rows']}@{q['w']['voxel'] * 1000:g}mm",
                          **t) for q, t in zip(seqs, tuning)],
        "autotune_what": "sicp_system_autotune after each warm-up round but the first: the "
                         "raycast's first-pass samples and second-pass workgroups with the lowest "
                         "summed median raycast time (raycast_us its mean per call, default_us "
                         "that of 48/1024); applied to every pass",
    }
    print(json.dumps(out), flush=True)


# The differentiable-renderer workload: 1280x720 / 2 mm voxels.
ADJOINT = dict(cols=1280, rows=720, f=1050.0, cx=639.5, cy=359.5, voxel=0.002, trunc=0.006,
               buckets=1 << 21, blocks=1_000_000, seed=0)


def run_adjoint(args):
    """The raycast adjoint at 1280x720 / 2 mm: 100 frames tracked and fused
    by the system, the model rendered at the last pose, then the backward pass
    of 10k seeded pixels with a hit, timed over `steps` calls after `warmup`.
    Also the agreement of each pixel's point-to-plane pose derivative
    n . dv/dxi with the ICP Jacobian [-n, n x v] (cosine; the march's two
    half-steps scale it by about 3/4 on surfaces seen head-on)."""
    import torch
    from semantic_icpslam_amd import binding as B
    w = ADJOINT
    nfr = 100
    log(f"[bench] rendering {nfr} frames of {w['cols']}x{w['rows']}")
    _, frames = make_frames(nfr, w)
    dev = torch.device("cuda:0")
    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    sysg = B.System(cfg)
    log("[bench] tracking and fusing the sequence")
    for d, c, l in frames:
        sysg.grab(torch.from_numpy(d).to(dev), torch.from_numpy(c).to(dev),
                  torch.from_numpy(l).to(dev))
    res = sysg.result()
    T = res["pose"]
    pose = B.pose_f(T)
    K = B.intrinsics(w["f"], w["f"], w["cx"], w["cy"])
    invf = float(np.float32(1.0 / w["f"]))
    zc, zr = (w["cols"] + 7) // 8, (w["rows"] + 7) // 8
    zmin = torch.zeros((zr, zc), dtype=torch.float32, device=dev)
    zmax = torch.zeros_like(zmin)
    nonempty = B.C.c_int(0)
    m = sysg.map_handle()
    B.check(B.lib().sicp_create_rendering_blocks(m, B.image(zmin), B.image(zmax), 3.0, 0.1,
                                                 B.C.byref(pose), None, K, B.C.byref(nonempty),
                                                 B.stream_ptr()))
    vmap = torch.empty((w["rows"], w["cols"], 4), dtype=torch.float32, device=dev)
    nmap = torch.empty_like(vmap)
    B.check(B.lib().sicp_raycast_inv(m, B.image(vmap), B.image(nmap), B.image(zmin),
                                     B.image(zmax), B.C.byref(pose), None, invf, invf, w["cx"],
                                     w["cy"], B.stream_ptr()))
    torch.cuda.synchronize()
    vh, nh = vmap.cpu().numpy(), nmap.cpu().numpy()
    ys, xs = np.nonzero(~np.isnan(vh[..., 0]))
    rng = np.random.default_rng(12345)
    sel = rng.choice(len(ys), 10000, replace=False)
    pix_h = (ys[sel] * w["cols"] + xs[sel]).astype(np.int32)
    pix = torch.from_numpy(pix_h).to(dev)
    n = len(pix_h)
    gv = torch.from_numpy(rng.normal(size=(n, 3)).astype(np.float32)).to(dev)
    gn = torch.from_numpy(rng.normal(size=(n, 3)).astype(np.float32)).to(dev)
    nblk = cfg.map.num_blocks
    gsdf = torch.zeros(nblk * 512, dtype=torch.float32, device=dev)
    gpx = torch.zeros((n, 6), dtype=torch.float32, device=dev)
    gpose = torch.zeros(6, dtype=torch.float64, device=dev)
    scratch = torch.zeros(B.lib().sicp_raycast_backward_scratch_size(), dtype=torch.uint8,
                          device=dev)

    def backward(gv_t, gn_t, sdf_t):
        B.check(B.lib().sicp_raycast_backward(
            m, B.image(zmin), B.image(zmax), B.C.byref(pose), invf, invf, w["cx"], w["cy"],
            w["cols"], w["rows"], pix.data_ptr(), n, gv_t.data_ptr(),
            gn_t.data_ptr() if gn_t is not None else None,
            sdf_t.data_ptr() if sdf_t is not None else None, gpx.data_ptr(), gpose.data_ptr(),
            scratch.data_ptr(), scratch.numel(), B.stream_ptr()))

    for _ in range(args.warmup):
        backward(gv, gn, gsdf)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        backward(gv, gn, gsdf)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    touched = int(torch.count_nonzero(gsdf).item())
    # point-to-plane pose derivative against the ICP Jacobian
    nn = torch.from_numpy(nh[ys[sel], xs[sel], :3].copy()).to(dev)
    backward(nn, None, None)
    torch.cuda.synchronize()
    a = gpx.cpu().numpy().astype(np.float64)
    nv = nh[ys[sel], xs[sel], :3].astype(np.float64)
    vv = vh[ys[sel], xs[sel], :3].astype(np.float64)
    J = np.concatenate([-nv, np.cross(nv, vv)], 1)
    cos = (a * J).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(J, axis=1) + 1e-30)
    ratio = (a * J).sum(1) / ((J * J).sum(1) + 1e-30)
    out = {
        "metric": "ms per raycast-adjoint call, 10k pixels, 1280x720 / 2 mm voxels",
        "value": round(ms, 4),
        "unit": "ms/call",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural room, seed 0, 100 frames fused)",
        "config": {"workload": "adjoint720", "pixels": n, "frames_fused": nfr,
                   "allocated_blocks": B.map_stats(m)["allocated_blocks"]},
        "pixels_per_s": round(n / (ms * 1e-3), 1),
        "touched_voxels": touched,
        "last_frame_tracked": res["tracked"],
        "icp_jacobian_agreement": {
            "what": "per pixel, n . dv/dxi from the adjoint against [-n, n x v]",
            "cosine_median": round(float(np.median(cos)), 6),
            "cosine_above_0_99": round(float(np.mean(cos > 0.99)), 4),
            "scale_median": round(float(np.median(ratio)), 4)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=60.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB first (counter calibration)")
    ap.add_argument("--no-pose-check", action="store_true")
    ap.add_argument("--no-photometric", action="store_true",
                    help="skip the pass with the photometric tracking term")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph-form pass (counter collection runs)")
    ap.add_argument("--no-scale", action="store_true",
                    help="skip the configured-scale passes (counter runs of the 5 mm workload)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the raycast's default launch shape (counter runs)")
    ap.add_argument("--ray-lds-table", type=int, default=1, choices=(0, 1),
                    help="0: the raycast's first pass without its LDS block table (counter "
                         "runs that isolate the hash probes' L2 traffic)")
    ap.add_argument("--workload", default="vga", choices=("vga", "batched8", "adjoint720"),
                    help="vga: the 640x480 / 5 mm sequence (default); batched8: eight "
                         "independent sequences at once; adjoint720: the raycast adjoint at "
                         "1280x720 / 2 mm")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent sequences would shard one per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched8":
        if args.steps == 290:  # the default frame count is that of the 640x480 sequence
            args.steps = 190
        run_batched(args)
        return
    if args.workload == "adjoint720":
        if args.steps == 290:
            args.steps = 50
        run_adjoint(args)
        return

    import torch
    from semantic_icpslam_amd import binding as B

    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {args.copy_calibration} MiB device copy")
        del src, dst

    w = WORKLOAD
    dev = torch.device("cuda:0")
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        sys.exit(2)
    # The whole sequence every pass: the untimed prefix (at least --warmup
    # frames) builds the map, the last --steps frames are timed.
    n = max(SEQ_FRAMES, args.warmup + args.steps)
    first_timed = n - args.steps
    log(f"[bench] rendering {n} frames of {w['cols']}x{w['rows']}")
    syn, frames = make_frames(n, w)
    depth = [torch.from_numpy(d).to(dev) for d, _, _ in frames]
    rgb = [torch.from_numpy(c).to(dev) for _, c, _ in frames]
    lab = [torch.from_numpy(lb).to(dev) for _, _, lb in frames]
    torch.cuda.synchronize()

    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          w["trunc"], w["buckets"], w["blocks"])
    stream = torch.cuda.current_stream()

    # The autotuner on a separate pass over the sequence, called after every
    # 30th frame; its sums over those frames pick the raycast launch shape.
    tune_frames = [] if args.no_autotune else list(range(29, n, 30))
    log(f"[bench] autotuning the raycast launch shape at {len(tune_frames)} frames")
    tuner = B.System(cfg)
    tuner.set_tuning(48, 1024, args.ray_lds_table)
    tuning, used = None, []
    for i in range(tune_frames[-1] + 1 if tune_frames else 0):
        tuner.grab(depth[i], rgb[i], lab[i], stream)
        if i in tune_frames:
            try:
                tuning = tuner.autotune(repeats=3, stream=stream)
                used.append(i)
            except B.SicpError:  # the frame was lost: nothing to time
                pass
    tuning = dict(tuning or tuner.tuning(), frames=used)
    tuner.close()
    log("[bench] tuning: " + json.dumps(tuning))

    def run_pass(timing, graph=False, system_cfg=None):
        """The whole sequence on a fresh system with the tuned raycast launch
        shape: frames before `first_timed` untimed, the rest timed; with
        `timing`, HIP events bracket every stage and kernel launch of the
        timed frames; with `graph`, every frame is copied into the system's
        staged inputs and runs as one hipGraph launch."""
        sysg = B.System(system_cfg or cfg)
        sysg.set_tuning(tuning["ray_first_pass_samples"], tuning["ray_second_pass_workgroups"],
                        tuning["ray_lds_block_table"])

        def grab(i):
            if graph:
                sysg.grab_staged(depth[i], rgb[i], lab[i], True, stream)
            else:
                sysg.grab(depth[i], rgb[i], lab[i], stream)

        for i in range(first_timed):
            grab(i)
        torch.cuda.synchronize()
        map_h = sysg.map_handle()
        st0 = B.map_stats(map_h, stream)
        if timing:
            B.check(B.lib().sicp_system_set_timing(sysg.h, 1))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(first_timed, n):
            grab(i)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        out = {"elapsed": elapsed, "st0": st0, "st1": B.map_stats(map_h, stream),
               "res": sysg.result(stream)}
        if timing:
            tot = np.zeros(len(STAGES))
            cnt = np.zeros(len(STAGES), np.int32)
            frame_ms = np.zeros(args.steps)
            nf = B.C.c_int(0)
            B.check(B.lib().sicp_system_read_timing_detail(sysg.h, STAGES.index("frame"),
                                                           frame_ms.ctypes.data, args.steps,
                                                           B.C.byref(nf), B.stream_ptr(stream)))
            B.check(B.lib().sicp_system_read_timing(sysg.h, tot.ctypes.data, cnt.ctypes.data,
                                                    len(STAGES), B.stream_ptr(stream)))
            B.check(B.lib().sicp_system_set_timing(sysg.h, 0))
            out.update(tot=tot, cnt=cnt, frame_ms=frame_ms[: nf.value])
        sysg.close()
        return out

    # Timed region: the last K frames with no instrumentation (the reported
    # value); then the same sequence again on a fresh system with per-stage
    # HIP events (stage times, per-frame percentiles, per-kernel rooflines).
    log(f"[bench] {first_timed} untimed + {args.steps} timed frames per pass")
    clean = run_pass(False)
    graphed = None
    if not args.no_graph:
        log("[bench] graph pass (staged inputs, one hipGraph launch per frame)")
        graphed = run_pass(False, graph=True)
    # The precision form (half TSDF, weights capped at 128, FP32 compensated
    # Gauss-Newton sums) in the graph form.
    cfg16 = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                            w["trunc"], w["buckets"], w["blocks"], tsdf_half=1, max_weight=128)
    cfg16.icp_sums_fp32 = 1
    precise = None
    if not args.no_graph:
        log("[bench] precision form pass (half TSDF, FP32 compensated sums, one hipGraph "
            "launch per frame)")
        precise = run_pass(False, graph=True, system_cfg=cfg16)
    photometric = None
    if not args.no_photometric:
        log("[bench] photometric pass (ComputeSE3(false): RGBStep + ICPStep per iteration)")
        cfg_rgb = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"],
                                  w["voxel"], w["trunc"], w["buckets"], w["blocks"])
        cfg_rgb.icp_only = 0
        photometric = run_pass(False, system_cfg=cfg_rgb)
    log("[bench] instrumented pass (HIP events per stage and kernel)")
    inst = run_pass(True)
    # The map at the configured heap size: the same frames at 2 mm voxels fill
    # about half of the 400k-block heap. Truncation 10 mm: at 6 mm (three
    # voxels) the reference's tracking loses this sequence from frame 1 on, as
    # the oracle does too.
    cfg_scale = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], 0.002,
                                0.010, w["buckets"], w["blocks"])
    scale_clean = scale_inst = None
    if not args.no_scale:
        log("[bench] configured-scale passes (2 mm voxels, the same frames)")
        scale_clean = run_pass(False, system_cfg=cfg_scale)
        scale_inst = run_pass(True, system_cfg=cfg_scale)
    st0, st1, res = inst["st0"], inst["st1"], clean["res"]
    tot, cnt, fm = inst["tot"], inst["cnt"], inst["frame_ms"]
    gt = syn.pose(n - 1)
    ms_per_frame = 1000.0 * clean["elapsed"] / args.steps
    stage_ms = {s_: round(float(tot[k] / max(cnt[k], 1)), 4) for k, s_ in enumerate(STAGES)}
    stage_ms["icp_iteration"] = round(float(tot[STAGES.index("icp_iteration")] / args.steps), 4)
    log("[bench] per-stage ms/frame: " + json.dumps(stage_ms))

    traffic = load_traffic()
    kernels = {}
    # Fusion: algorithmic bytes of the timed frames (the map's own counters)
    # over the summed fuse-kernel event time.
    upd = st1["updated_total"] - st0["updated_total"]
    vis = st1["visible_total"] - st0["visible_total"]
    heap = st1["heap_scanned_total"] - st0["heap_scanned_total"]
    fus = max(st1["fusions"] - st0["fusions"], 1)
    fuse_bytes = upd * FUSE_BYTES_PER_VOXEL + vis * FUSE_BYTES_PER_BLOCK
    kernels["fuse_kernel"] = roofline_entry("fuse_kernel", fuse_bytes,
                                            tot[STAGES.index("fuse")] / 1000.0, fus,
                                            traffic.get("fuse_kernel"))
    kernels["fuse_kernel"]["updated_voxels_per_launch"] = int(upd / fus)
    kernels["fuse_kernel"]["visible_blocks_per_launch"] = int(vis / fus)

    cpu, work = None, None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle)")
        cpu, work = cpu_baseline(w, frames, first_timed, args.cpu_seconds)
    ray_bytes = alloc_bytes = None
    if work and work["raycasts"]:
        # Raycast and allocation: the oracle's counts over the same frames
        # (deterministic; scaled by frames when the CPU leg covered only a
        # prefix of the timed frames).
        scale = args.steps / work["frames"]
        ray_bytes = scale * (RAY_BYTES_PER_RAY * work["rays"] +
                             RAY_BYTES_PER_STEP * work["march_steps"] +
                             RAY_BYTES_PER_INTERP * work["interp_reads"] +
                             RAY_BYTES_PER_LOOKUP * work["ray_lookups"])
        alloc_bytes = (scale * ALLOC_BYTES_PER_PROBE * work["alloc_probes"] +
                       args.steps * ALLOC_BYTES_PER_PIXEL * w["cols"] * w["rows"])
        ray_traffic = None
        if "raycast_kernel" in traffic and "raycast_long_kernel" in traffic:
            ray_traffic = traffic["raycast_kernel"] + traffic["raycast_long_kernel"]
        kernels["raycast"] = roofline_entry(
            "raycast_kernel + raycast_long_kernel", ray_bytes,
            tot[STAGES.index("raycast")] / 1000.0, cnt[STAGES.index("raycast")], ray_traffic)
        kernels["raycast"]["work_per_ray"] = {
            k: round(work[k] / max(work["rays"], 1), 3)
            for k in ("march_steps", "interp_reads", "ray_lookups")}
        kernels["raycast"]["work_frames"] = work["frames"]
        if ray_traffic:
            # the same launches' measured HBM traffic (KiB per launch) over their time
            kernels["raycast"]["traffic_gbs"] = round(
                ray_traffic * 1024 * cnt[STAGES.index("raycast")] /
                max(tot[STAGES.index("raycast")] / 1000.0, 1e-12) / 1e9, 1)
    # ICP: every Gauss-Newton launch reads 64 B per pixel of its level.
    lv = [(w["cols"] >> l_) * (w["rows"] >> l_) for l_ in range(3)]
    icp_launches = int(cnt[STAGES.index("icp_iteration")])
    icp_bytes = ICP_BYTES_PER_PIXEL * args.steps * (10 * lv[0] + 5 * lv[1] + 3 * lv[2])
    kernels["icp_iteration_kernel"] = roofline_entry(
        "icp_iteration_kernel", icp_bytes, tot[STAGES.index("icp_iteration")] / 1000.0,
        icp_launches, traffic.get("icp_iteration_kernel"))
    kernels["icp_iteration_kernel"]["launches_per_frame"] = round(icp_launches / args.steps, 2)
    # The dominant kernel: the most time per frame (summed over its launches).
    dom = max(kernels, key=lambda k: kernels[k]["avg_launch_us"] *
              kernels[k].get("launches_per_frame", 1.0))
    roofline = dict(kernels[dom])
    roofline["traffic_source"] = (
        "profiles/pmc_traffic_vga_5mm_sequence.json: rocprofv3 FETCH_SIZE and WRITE_SIZE "
        "passes over bench.py --no-graph --no-autotune --no-scale on this 300-frame workload "
        "(the same kernels, default raycast shape), each scaled by the same counter on a plain "
        "device copy")

    # Every stage of the frame: algorithmic bytes per frame over the stage's
    # event time (the instrumented pass, timed frames).
    zc, zr = (w["cols"] + 7) // 8, (w["rows"] + 7) // 8
    stage_bytes = {
        "pyramid": args.steps * pyramid_bytes(w["cols"], w["rows"]),
        "track": icp_bytes,
        "allocate": alloc_bytes,
        "visibility": VIS_BYTES_PER_SLOT * heap + VIS_BYTES_PER_VISIBLE * vis,
        "fuse": fuse_bytes,
        "prepass": PREPASS_BYTES_PER_BLOCK * vis + args.steps * 8 * zc * zr,
        "raycast": ray_bytes,
    }
    stages = {}
    for name, nbytes in stage_bytes.items():
        ms = tot[STAGES.index(name)] / args.steps
        e = {"ms_per_frame": round(float(ms), 4)}
        if nbytes:
            gbs = nbytes / args.steps / (ms * 1e-3) / 1e9
            e.update(bytes_per_frame=int(nbytes / args.steps), gbs=round(gbs, 1),
                     frac=round(gbs / HBM_PEAK_GBS, 4))
        stages[name] = e
    known = [b for b in stage_bytes.values() if b]
    if len(known) == len(stage_bytes):
        fr_ms = tot[STAGES.index("frame")] / args.steps
        gbs = sum(known) / args.steps / (fr_ms * 1e-3) / 1e9
        stages["frame"] = {"ms_per_frame": round(float(fr_ms), 4),
                           "bytes_per_frame": int(sum(known) / args.steps),
                           "gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    stages["what"] = ("algorithmic bytes per frame (each input read once, each output written "
                      "once) over the stage's HIP-event time in the instrumented pass; "
                      "allocation and raycast counts from the oracle on the same frames; the "
                      "track stage's time includes its finishing launch")

    small = None
    if not args.no_cpu_baseline:
        log("[bench] CPU reference method at the small size")
        small = cpu_small()
    pose_check = prec = None
    if not args.no_pose_check:
        log("[bench] pose precision against FP64 (both forms, whole sequence)")
        pose_check = pose_precision(w, frames, syn, {"fp32_form": cfg, "precision_form": cfg16})
        log("[bench] TSDF and depth of both forms against FP64")
        prec = precision_report(w, frames, syn, cfg, cfg16, m=min(30, n))

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
                   "frames": n, "timed_frames": [first_timed, n - 1], "global_batch": 1},
        "stage_ms_per_frame": stage_ms,
        "stages": stages,
        "autotune": dict(tuning, what="sicp_system_autotune on a separate pass, after each "
                         "of these frames: the raycast launch shape with the lowest summed "
                         "median time (raycast_us its mean per call, default_us that of 48 "
                         "samples / 1024 workgroups), used by every pass; any shape gives "
                         "bit-identical maps"),
        "allocated_blocks": st1["allocated_blocks"],
        "last_frame_tracked": res["tracked"],
        "last_frame_translation_vs_ground_truth_m": round(
            float(np.linalg.norm(res["pose"][:3, 3] - gt[:3, 3])), 5),
        "graph_form": None if graphed is None else {
            "ms_per_step": round(1000.0 * graphed["elapsed"] / args.steps, 4),
            "what": "the same frames, each copied into the system's staged inputs and run as "
                    "one hipGraph launch (plain FP32)",
            "same_final_pose": bool(np.array_equal(graphed["res"]["pose"], res["pose"]))},
        "precision_form": None if precise is None else {
            "ms_per_step": round(1000.0 * precise["elapsed"] / args.steps, 4),
            "what": "the same frames with the sdf stored as IEEE half (one 4-byte word with the "
                    "weight; colour in the semantic record), weights capped at 128, FP32 "
                    "updates and interpolation, the ICP sums in FP32 with compensated "
                    "summation (TwoSum at every addition; FP64 only in the 6x6 solve and the "
                    "pose composition), one hipGraph launch per frame",
            "voxel_bytes": {"fp32_form": 8 + 12, "half_form": 4 + 12},
            "allocated_blocks": precise["st1"]["allocated_blocks"],
            "last_frame_tracked": precise["res"]["tracked"],
            "last_frame_translation_vs_ground_truth_m": round(
                float(np.linalg.norm(precise["res"]["pose"][:3, 3] - gt[:3, 3])), 5),
            "vs_fp64": prec},
        "pose_precision": pose_check,
        "at_scale": None if scale_inst is None else {
            "what": "the same frames with 2 mm voxels and 10 mm truncation (2^20 x 2 buckets, "
                    "400k-block heap), plain FP32, eager launches: the heap-slot visibility "
                    "scan and the table at the configured heap scale",
            "ms_per_step": round(1000.0 * scale_clean["elapsed"] / args.steps, 4),
            "allocated_blocks": scale_inst["st1"]["allocated_blocks"],
            "heap_blocks": w["blocks"],
            "visible_blocks_per_frame": int(
                (scale_inst["st1"]["visible_total"] - scale_inst["st0"]["visible_total"]) /
                max(scale_inst["st1"]["fusions"] - scale_inst["st0"]["fusions"], 1)),
            "stage_ms_per_frame": {
                s_: round(float(scale_inst["tot"][k] / max(scale_inst["cnt"][k], 1)), 4)
                for k, s_ in enumerate(STAGES) if s_ not in ("icp_iteration", "unify")},
            "last_frame_tracked": scale_clean["res"]["tracked"],
            "last_frame_translation_vs_ground_truth_m": round(
                float(np.linalg.norm(scale_clean["res"]["pose"][:3, 3] - gt[:3, 3])), 5)},
        "photometric_tracking": None if photometric is None else {
            "ms_per_step": round(1000.0 * photometric["elapsed"] / args.steps, 4),
            "what": "the same frames with the reference's default tracking, ComputeSE3(false): "
                    "RGBStep (correspondence pass fused into the ICP pass, then a row pass) "
                    "and ICPStep in every iteration, photometric weight 1e-4",
            "last_frame_tracked": photometric["res"]["tracked"],
            "last_frame_translation_vs_ground_truth_m": round(
                float(np.linalg.norm(photometric["res"]["pose"][:3, 3] - gt[:3, 3])), 5)},
        "frame_ms_p50": round(float(np.percentile(fm, 50)), 4) if fm.size else None,
        "frame_ms_p99": round(float(np.percentile(fm, 99)), 4) if fm.size else None,
        "roofline": roofline,
        "roofline_kernels": kernels,
        "cpu_baseline": cpu,
        "cpu_small": small,
    }
    print(json.dumps(out), flush=True)

if __name__ == "__main__":
    main()
