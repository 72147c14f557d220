"""The raycast's launch shape (first-pass samples before the hand-over,
second-pass workgroups) and the autotuner that picks it: every shape gives
the same model maps, poses and map to the bit, so tuning changes time only.
A changed tuning re-records the system's graphs (seen in the hand-over
count, which depends on the first-pass samples)."""
import numpy as np
import pytest
import torch

from helpers import B, CFG_VGA, download, sys_cfg

pytestmark = pytest.mark.gpu

FRAMES = 8
# (first-pass samples, second-pass workgroups, LDS block table)
SHAPES = [(48, 1024, 1), (4, 64, 1), (4096, 1, 1), (17, 300, 1), (48, 1024, 0), (4, 64, 0)]


@pytest.fixture(scope="module")
def frames(dev):
    c = CFG_VGA
    s = B.Synth(c["cols"], c["rows"], c["f"], c["f"], c["cx"], c["cy"], seed=11)
    out = [tuple(torch.from_numpy(a).to(dev) for a in s.render(i)) for i in range(FRAMES)]
    torch.cuda.synchronize()
    return out


def _model(g):
    c = CFG_VGA
    v = download(g.frame_image(0, 0), c["rows"], c["cols"], np.float32, 4)
    n = download(g.frame_image(1, 0), c["rows"], c["cols"], np.float32, 4)
    return v, n


def _run(frames, shape=None, graph=False, autotune_at=()):
    g = B.System(sys_cfg(CFG_VGA))
    if shape:
        g.set_tuning(*shape)
    res, handover, tuned = [], [], []
    for i, (d, c, lab) in enumerate(frames):
        if graph:
            g.grab_staged(d, c, lab, True)
        else:
            g.grab(d, c, lab)
        r = g.result()
        res.append((r, _model(g)))
        handover.append(B.map_stats(g.map_handle())["ray_handover"])
        if i in autotune_at:
            tuned.append((g.autotune(repeats=3), g.autotune_results()))
    torch.cuda.synchronize()
    m = B.Map.__new__(B.Map)
    m.cfg, m.h = g.cfg.map, g.map_handle()
    d = m.download(n_blocks=B.map_stats(m.h)["allocated_blocks"])
    m.h = None
    g.close()
    return res, d, handover, tuned


@pytest.fixture(scope="module")
def default_run(frames):
    return _run(frames)


def _same(got, want):
    (rg, dg, _, _), (rw, dw, _, _) = got, want
    for i, ((a, (va, na)), (b, (vb, nb))) in enumerate(zip(rg, rw)):
        assert np.array_equal(a["pose"], b["pose"]), i
        assert a["tracked"] == b["tracked"] and a["icp_count"] == b["icp_count"], i
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), i
        assert np.array_equal(na.view(np.uint32), nb.view(np.uint32)), i
    for f in ("table", "excess", "blocks", "sem"):
        assert np.array_equal(dg[f], dw[f]), f
    assert np.array_equal(dg["geom"].view(np.uint8), dw["geom"].view(np.uint8))


@pytest.mark.parametrize("shape", SHAPES[1:])
def test_launch_shapes_equal_default(frames, default_run, shape):
    print("shape", shape, flush=True)
    got = _run(frames, shape)
    _same(got, default_run)
    tracked = sum(r["tracked"] for r, _ in default_run[0][1:])
    assert tracked >= FRAMES - 2, tracked
    # fewer first-pass samples hand more rays over
    if shape[0] < 48:
        assert sum(got[2]) > sum(default_run[2]) > 0
    if shape[0] == 4096:
        assert sum(got[2]) == 0


def test_autotune_in_graph_form_equals_default(frames, default_run):
    """Autotuned after frames 2 and 3 (sums over both calls), in the graph
    form: the maps stay equal to the default shape's."""
    got = _run(frames, (4, 64), graph=True, autotune_at=(2, 3))
    _same(got, default_run)
    (first, table1), (tuned, table) = got[3]
    print("autotuned", tuned, flush=True)
    assert len(table) == 28 and table[0][:2] == (48, 1024)
    assert (tuned["ray_first_pass_samples"], tuned["ray_second_pass_workgroups"]) in \
        [t[:2] for t in table]
    means = [t[2] for t in table]
    assert 0 < tuned["raycast_us"] == min(means) and tuned["default_us"] == means[0]
    # the second call's means average both calls
    assert means != [t[2] for t in table1]
    assert tuned["ray_lds_block_table"] == 1
    # the graphs were recorded again with the new first-pass samples (never 4)
    handover_4 = _run(frames, (4, 64))[2]
    assert got[2][:3] == handover_4[:3]
    assert got[2][3:] != handover_4[3:]


def test_autotune_keeps_the_block_table_setting(frames, default_run):
    got = _run(frames, (4, 64, 0), autotune_at=(2,))
    _same(got, default_run)
    (tuned, table), = got[3]
    assert tuned["ray_lds_block_table"] == 0 and len(table) == 28


def test_block_table_change_clears_the_autotuner_sums(frames):
    g = B.System(sys_cfg(CFG_VGA))
    for d, c, lab in frames[:3]:
        g.grab(d, c, lab)
    g.autotune(repeats=1)
    assert len(g.autotune_results()) == 28
    g.set_tuning(32, 512, 1)  # same table setting: the sums stay
    assert len(g.autotune_results()) == 28
    g.set_tuning(32, 512, 0)
    assert g.autotune_results() == []
    g.close()


def test_tuning_arguments(dev):
    g = B.System(sys_cfg(CFG_VGA))
    assert g.tuning() == dict(ray_first_pass_samples=48, ray_second_pass_workgroups=1024,
                              ray_lds_block_table=1)
    for bad in ((0, 1024, 1), (4097, 1024, 1), (48, 0, 1), (48, 65537, 1), (48, 1024, 2)):
        t = B.Tuning(*bad)
        assert B.lib().sicp_system_set_tuning(g.h, B.C.byref(t)) == B.SICP_ERR_INVALID_ARGUMENT
    assert g.tuning()["ray_first_pass_samples"] == 48
    # nothing to raycast before the first frame
    t = B.Tuning()
    assert B.lib().sicp_system_autotune(g.h, 3, B.C.byref(t), None, None, None) == \
        B.SICP_ERR_INVALID_ARGUMENT
    assert B.lib().sicp_system_autotune(g.h, 0, B.C.byref(t), None, None, None) == \
        B.SICP_ERR_INVALID_ARGUMENT
    g.close()
