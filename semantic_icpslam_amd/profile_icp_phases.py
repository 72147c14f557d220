"""Per-phase timing of the tracking loop's launches (icp_iteration_kernel) on
the seeded 640x480 / 5 mm room sequence, from the profiling build's
wall-clock stamps (100 MHz).

    make -C semantic_icpslam_amd/csrc phases
    python semantic_icpslam_amd/profile_icp_phases.py [--frames 40] [--skip 10]
        [--precision-form]

Prints one row per launch of the frame's 18 (levels 2, 1, 0 with 3, 5 and 10
iterations), medians over the frames after `skip`:
  span      first workgroup's start to last workgroup's end
  fold      start to the previous iteration's partial rows summed (this
            includes the first loads' latency)
  gn        the Gauss-Newton update on one wavefront, of which
  solve     the 6x6 LDL^T solve (with the error tests and A, b set-up)
  exp       the SE(3) exponential
  publish   state ready in LDS (and published by workgroup 0)
  pixel     the pixel pass (correspondences, 29 products per pixel)
  reduce    the workgroup's LDS reduction and its partial-row store
  spread    last workgroup start minus first (dispatch)
  gap       this launch's last end to the next launch's first start
Per-workgroup phases are medians over the launch's workgroups.
"""
import argparse
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PHASE_LIB = os.path.join(HERE, "libsicp_phases.so")
LAUNCHES, MAX_WG, STAMPS = 24, 512, 8  # the profiling build's stamp array


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--skip", type=int, default=10)
    ap.add_argument("--json", default=None, help="also write the rows here")
    ap.add_argument("--precision-form", action="store_true",
                    help="half TSDF, weights to 128, FP32 compensated sums")
    args = ap.parse_args()
    if not os.path.exists(PHASE_LIB):
        sys.exit(f"{PHASE_LIB} missing: run make -C semantic_icpslam_amd/csrc phases")
    os.environ["SICP_LIB"] = PHASE_LIB
    sys.path.insert(0, os.path.dirname(HERE))
    import torch
    from semantic_icpslam_amd import binding as B
    L = B.lib()
    L.sicp_profile_icp_phases.restype = ctypes.c_int
    L.sicp_profile_icp_phases.argtypes = [ctypes.c_void_p]

    cols, rows, f = 640, 480, 525.0
    syn = B.Synth(cols, rows, f, f, 319.5, 239.5, seed=0)
    cfg = B.system_config(cols, rows, f, f, 319.5, 239.5, 0.005, 0.015, 1 << 20, 400000)
    if args.precision_form:
        cfg.map.tsdf_half, cfg.map.max_weight, cfg.icp_sums_fp32 = 1, 128, 1
    iters = [cfg.icp_iters[2], cfg.icp_iters[1], cfg.icp_iters[0]]
    levels = [2] * iters[0] + [1] * iters[1] + [0] * iters[2]
    def workgroups(n):  # the launch's grid: 4 pixels per thread, 1 up to 320 * 256 pixels
        return min(MAX_WG, -(-n // (256 if n <= 320 * 256 else 1024)))
    nwg = [workgroups((cols >> l) * (rows >> l)) for l in levels]
    dev = torch.device("cuda:0")
    g = B.System(cfg)
    buf = np.zeros((LAUNCHES, MAX_WG, STAMPS), np.uint64)
    per = [[] for _ in levels]
    gaps = [[] for _ in levels]
    for i in range(args.frames):
        d, c, lab = (torch.from_numpy(a).to(dev) for a in syn.render(i))
        g.grab(d, c, lab)
        torch.cuda.synchronize()
        if i < args.skip:
            continue
        r = g.result()
        if not r["tracked"] or r["iterations"] != len(levels):
            continue  # a launch ran no update: its stamps are partial
        B.check(L.sicp_profile_icp_phases(buf.ctypes.data), "profile_icp_phases")
        t = buf.astype(np.int64)
        starts, ends = [], []
        for k, n in enumerate(nwg):
            a = t[k, :n]
            starts.append(a[:, 0].min())
            ends.append(a[:, 5].max())
            ph = np.median(np.diff(a[:, :6], axis=1), axis=0)  # fold gn publish pixel reduce
            sub = (np.median(a[:, 6] - a[:, 1]), np.median(a[:, 7] - a[:, 6])) if k else (0, 0)
            per[k].append([a[:, 5].max() - a[:, 0].min(), *ph, a[:, 0].max() - a[:, 0].min(),
                           *sub])
        for k in range(len(levels) - 1):
            gaps[k].append(starts[k + 1] - ends[k])
        print(f"frame {i} stamped", file=sys.stderr, flush=True)
    g.close()
    if not per[0]:
        sys.exit("no frame ran all launches")
    out = []
    print("| launch | level | workgroups | span | fold | gn | publish | pixel | reduce | "
          "spread | gap | solve | exp |")
    print("|---|---|---|---|---|---|---|---|---|---|---|---|---|")
    for k, (l, n) in enumerate(zip(levels, nwg)):
        m = np.median(np.array(per[k]), axis=0) / 100.0  # 100 MHz ticks -> us
        gap = float(np.median(gaps[k])) / 100.0 if gaps[k] else float("nan")
        row = dict(launch=k, level=l, workgroups=n, span=m[0], fold=m[1], gn=m[2],
                   publish=m[3], pixel=m[4], reduce=m[5], spread=m[6], gap=gap, solve=m[7],
                   exp=m[8])
        out.append({k_: (round(float(v), 2) if isinstance(v, (float, np.floating)) else v)
                    for k_, v in row.items()})
        print("| " + " | ".join(f"{v:.2f}" if isinstance(v, float) else str(v)
                                for v in out[-1].values()) + " |")
    spans = sum(r["span"] for r in out)
    gsum = sum(r["gap"] for r in out if r["gap"] == r["gap"])
    print(f"\nframes: {len(per[0])}; sum of spans {spans:.1f} us, sum of gaps {gsum:.1f} us "
          "(microseconds, medians)")
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(out, fh)


if __name__ == "__main__":
    main()
