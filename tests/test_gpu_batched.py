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
    """The 320x240 sequences keep one pixel per thread on their small levels
    (icp_spread, the default), the others four pixels everywhere, as the
    batched bench runs them: both partitions in one batch."""
    cfg = B.system_config(w["cols"], w["rows"], w["f"], w["f"], w["cx"], w["cy"], w["voxel"],
                          3 * w["voxel"], 1 << 18, 60000)
    cfg.icp_spread = 1 if w["cols"] == 320 else 0
    return cfg


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


@pytest.fixture(scope="module")
def alone(dev):
    """Each sequence run by itself: per-frame results and final map."""
    frames = [_frames(w, dev) for w in SEQS]
    torch.cuda.synchronize()
    out = []
    for k, w in enumerate(SEQS):
        g = B.System(_config(w))
        res = _run([g], [frames[k]], [torch.cuda.Stream()])[0]
        out.append((res, _map_state(g)))
        g.close()
    return frames, out


def _same(k, got, want):
    (res_g, (ng, dg)), (res_w, (nw, dw)) = got, want
    for i, (a, b) in enumerate(zip(res_g, res_w)):
        assert np.array_equal(a["pose"], b["pose"]), (k, i)
        for f in ("tracked", "icp_count", "iterations", "visible_blocks", "allocated_blocks"):
            assert a[f] == b[f], (k, i, f)
        assert a["icp_error"] == b["icp_error"] or (np.isnan(a["icp_error"]) and
                                                    np.isnan(b["icp_error"]))
    assert ng == nw and ng > 0
    for f in ("table", "excess", "blocks"):
        assert np.array_equal(dg[f], dw[f]), (k, f)
    assert np.array_equal(dg["geom"].view(np.uint8), dw["geom"].view(np.uint8)), k
    assert np.array_equal(dg["sem"], dw["sem"]), k


def test_eight_sequences_concurrent_equal_alone(alone):
    frames, want = alone
    systems = [B.System(_config(w)) for w in SEQS]
    streams = [torch.cuda.Stream() for _ in SEQS]
    together = _run(systems, frames, streams)
    got = [(together[k], _map_state(g)) for k, g in enumerate(systems)]
    for g in systems:
        g.close()
    tracked = 0
    for k in range(len(SEQS)):
        print(f"sequence {k}", flush=True)
        _same(k, got[k], want[k])
        tracked += sum(r["tracked"] for r in want[k][0][1:])
    assert tracked >= len(SEQS) * (FRAMES - 1) * 3 // 4, tracked


@pytest.mark.parametrize("graph,shared", [(False, True), (True, True), (True, False)])
def test_eight_sequences_one_launch_per_iteration_equal_alone(alone, graph, shared):
    """The eight sequences advanced as one batch: their tracking iterations
    share launches (sicp_system_grab_batch), eagerly or as one hipGraph per
    round (sicp_batch_grab_staged); or, unshared, each system's frame as an
    independent chain of the round's graph."""
    frames, want = alone
    systems = [B.System(_config(w)) for w in SEQS]
    batch = B.Batch(systems) if graph else None
    if batch and not shared:
        batch.set_shared_tracking(False)
    res = [[] for _ in SEQS]
    for i in range(FRAMES):
        fr = [frames[k][i] for k in range(len(SEQS))]
        if graph:
            batch.grab_staged(fr, True)
        else:
            B.grab_batch(systems, fr)
        for k, g in enumerate(systems):
            res[k].append(g.result())
    torch.cuda.synchronize()
    got = [(res[k], _map_state(g)) for k, g in enumerate(systems)]
    if batch:
        batch.close()
    for g in systems:
        g.close()
    for k in range(len(SEQS)):
        print(f"sequence {k} ({'graph' if graph else 'eager'}, shared {shared})", flush=True)
        _same(k, got[k], want[k])


def test_batch_rejects_mixed_forms(dev):
    w = SEQS[3]
    a = B.System(_config(w))
    cfg = _config(w)
    cfg.icp_only = 0
    b = B.System(cfg)
    hs = (B.C.c_void_p * 2)(a.h, b.h)
    h = B.C.c_void_p()
    assert B.lib().sicp_batch_create(hs, 2, B.C.byref(h)) == B.SICP_ERR_INVALID_ARGUMENT
    hs = (B.C.c_void_p * 2)(a.h, a.h)
    assert B.lib().sicp_batch_create(hs, 2, B.C.byref(h)) == B.SICP_ERR_INVALID_ARGUMENT
