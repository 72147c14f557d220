"""Eight independent sequences of uneven resolution and voxel size
(3 x 640x480 at 5 mm, 3 x 320x240 at 8 mm, 2 x 1280x720 at 4 mm), each with
its own system, run together: one stream per sequence, every frame after the
first one hipGraph launch. Each sequence's per-frame poses, tracking results
and final map must equal those of the same sequence run alone, to the bit:
the kernels share nothing between systems and every reduction has a fixed
order, so concurrency may change timing only."""
import numpy as np
import pytest
import torch

from helpers import B

pytestmark = pytest.mark.gpu

SEQS = (
    [dict(cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.005, seed=k) for k in (0, 1, 2)] +
    [dict(cols=320, rows=240, f=262.5, cx=159.5, cy=119.5, voxel=0.008, seed=k) for k in (3, 4, 5)] +
    [dict(cols=1280, rows=720, f=1050.0, cx=639.5, cy=359.5, voxel=0.004, seed=k) for k in (6, 7)])
FRAMES = 5


def _config(w):
    return B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                           3 * w["voxel"], 1 << 18, 60000)


def _frames(w, dev):
    s = B.Synth(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], seed=w["seed"])
    return [tuple(torch.from_numpy(a).to(dev) for a in s.render(i)) for i in range(FRAMES)]


def _map_state(g):
    h = g.map_handle()
    st = B.map_stats(h)
    n = st["allocated_blocks"]
    m = B.Map.__new__(B.Map)  # a view of the system's map (not owned)
    m.cfg, m.h = g.cfg.map, h
    d = m.download(n_blocks=n)
    m.h = None
    return n, d


def _run(systems, frames, streams):
    """Advances every system through every frame, one round at a time, and
    returns each system's per-frame results."""
    out = [[] for _ in systems]
    for i in range(FRAMES):
        for g, fr, st in zip(systems, frames, streams):
            d, c, lab = fr[i]
            g.grab_staged(d, c, lab, True, st)
        for k, (g, st) in enumerate(zip(systems, streams)):
            out[k].append(g.result(st))
    torch.cuda.synchronize()
    return out


def test_eight_sequences_concurrent_equal_alone(dev):
    frames = [_frames(w, dev) for w in SEQS]
    tracked = 0
    torch.cuda.synchronize()
    systems = [B.System(_config(w)) for w in SEQS]
    streams = [torch.cuda.Stream() for _ in SEQS]
    together = _run(systems, frames, streams)
    maps_together = [_map_state(g) for g in systems]
    for g in systems:
        g.close()
    for k, w in enumerate(SEQS):
        print(f"sequence {k}: {w['cols']}x{w['rows']} seed {w['seed']}", flush=True)
        g = B.System(_config(w))
        alone = _run([g], [frames[k]], [torch.cuda.Stream()])[0]
        n, d = _map_state(g)
        g.close()
        for i, (a, b) in enumerate(zip(together[k], alone)):
            assert np.array_equal(a["pose"], b["pose"]), (k, i)
            for f in ("tracked", "icp_count", "iterations", "visible_blocks",
                      "allocated_blocks"):
                assert a[f] == b[f], (k, i, f)
            assert a["icp_error"] == b["icp_error"] or (np.isnan(a["icp_error"]) and
                                                        np.isnan(b["icp_error"]))
        nt, dt = maps_together[k]
        assert n == nt and n > 0
        for f in ("table", "excess", "blocks"):
            assert np.array_equal(d[f], dt[f]), (k, f)
        assert np.array_equal(d["geom"].view(np.uint8), dt["geom"].view(np.uint8)), k
        assert np.array_equal(d["sem"], dt["sem"]), k
        tracked += sum(r["tracked"] for r in alone[1:])
    assert tracked >= len(SEQS) * (FRAMES - 1) * 3 // 4, tracked
