"""The whole per-frame path (track -> fuse -> raycast) on the GPU against the
oracle's restatement of System::GrabImage on the same seeded frames.

The ICP sums are accumulated in a different order on the two sides, so the
poses agree to a tolerance, and the maps they build agree statistically:
per-frame pose difference below 1e-4 (translation in metres, rotation
entries), tracked flags equal, allocated-block counts within 1 %.
"""
import numpy as np
import pytest

from helpers import B, CFG_SMALL, CFG_VGA, sys_cfg, synth, to_dev
from oracle import binding as O

pytestmark = pytest.mark.gpu


# half = 1: the precision form's map (sdf as half, weights up to 128);
# k32 = 1: its FP32 compensated Gauss-Newton sums, on both sides; spread = 0:
# four pixels per thread at every level (the oracle follows the partition).
@pytest.mark.parametrize("c,frames,half,k32,spread", [
    (CFG_SMALL, 12, 0, 0, 1), (CFG_VGA, 4, 0, 0, 1), (CFG_SMALL, 8, 1, 0, 1),
    (CFG_SMALL, 8, 1, 1, 1), (CFG_VGA, 4, 1, 1, 1), (CFG_SMALL, 8, 1, 1, 0)])
def test_system_vs_oracle(dev, c, frames, half, k32, spread):
    cfg = sys_cfg(c, **(dict(tsdf_half=1, max_weight=128) if half else {}))
    cfg.icp_sums_fp32 = k32
    cfg.icp_spread = spread
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    s = synth(c)
    for i in range(frames):
        d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        rg = g.result()
        ro = o.grab(d, rgb, lab)
        assert rg["tracked"] == ro["tracked"], (i, rg, ro)
        dp = np.abs(rg["pose"] - ro["pose"]).max()
        assert dp < 1e-4, (i, dp)
        na, nb = rg["allocated_blocks"], ro["allocated_blocks"]
        assert abs(na - nb) <= 0.01 * nb + 2, (i, na, nb)
        if i == 0:
            assert na == nb and rg["visible_blocks"] == ro["visible_blocks"]


@pytest.mark.parametrize("jump,tracked", [(9, 0), (8, 1)])
def test_pose_jump_motion_gate(dev, jump, tracked):
    """Frame 0 then frame `jump` (true motion 0.104 m / 0.093 m): tracking
    converges either way; the 0.1 m translation gate (Tracking.cc:505-516)
    loses the first frame, restoring the pose and skipping its fusion, and
    keeps the second, as the oracle does."""
    c = CFG_SMALL
    cfg = sys_cfg(c)
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    s = synth(c)
    out = []
    for i in (0, jump):
        d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        out.append((g.result(), o.grab(d, rgb, lab)))
    (g0, _), (rg, ro) = out
    assert rg["tracked"] == ro["tracked"] == tracked, (rg, ro)
    assert 0 < rg["icp_error"] < cfg.icp_thresh
    assert rg["allocated_blocks"] == ro["allocated_blocks"]
    assert np.abs(rg["pose"] - ro["pose"]).max() < 1e-4
    if not tracked:
        assert rg["iterations"] == 18 and np.array_equal(rg["pose"], g0["pose"])
        assert rg["allocated_blocks"] == g0["allocated_blocks"]


def test_system_first_frame_model_bit_exact(dev):
    """Frame 0 has no tracking: its fused map and raycast model must match the
    oracle exactly."""
    c = CFG_SMALL
    cfg = sys_cfg(c)
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    d, rgb, lab = synth(c).render(0)
    g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
    o.grab(d, rgb, lab)
    import torch
    torch.cuda.synchronize()
    im = g.frame_image(0, 0)
    rows, cols = c["rows"], c["cols"]
    buf = np.zeros((rows, cols, 4), np.float32)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipMemcpy2D(buf.ctypes.data, cols * 16, im.ptr, im.pitch, cols * 16, rows, 2) == 0
    ov, _ = o.model(0)
    from helpers import bits_equal, mismatch_report
    assert bits_equal(buf, ov), mismatch_report(buf, ov)
    # The fusion's counters (summed by the render prepass in this path) equal
    # the oracle's fused-voxel and visible-block counts of the same frame.
    st, w = B.map_stats(g.map_handle()), o.work()
    assert st["fusions"] == 1
    assert st["updated_voxels"] == st["updated_total"] == w["updated_voxels"] > 0
    assert st["visible_total"] == w["visible_blocks"] > 0


@pytest.mark.parametrize("c", [CFG_SMALL, dict(CFG_SMALL, cols=317, rows=239, cx=158.0, cy=119.0),
                               CFG_VGA])
def test_system_frame_pyramid_bit_exact(dev, c):
    """The three-launch frame pyramid (filter, intensity, PyrDown to levels
    1-2, vertex / normal maps, derivatives) and the model pyramid of the
    previous frame (ResizeMap to levels 1-2) equal the single-image oracle
    functions applied level by level. The odd 317x239 size exercises the
    clipped windows and the odd remainder rows / columns."""
    from helpers import bits_equal, download, mismatch_report
    cfg = sys_cfg(c)
    g = B.System(cfg)
    s = synth(c)
    frames = [s.render(i) for i in range(2)]
    for d, rgb, lab in frames:
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
    d, rgb, _ = frames[1]
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    depth, gray = [filt], [O.image_to_intensity(rgb)]
    for _ in range(2):
        depth.append(O.pyrdown_f32(depth[-1]))
        gray.append(O.pyrdown_u8(gray[-1]))
    rows0, cols0 = c["rows"], c["cols"]
    got = download(g.frame_image(2, 0), rows0, cols0, np.float32)
    assert bits_equal(got, raw), mismatch_report(got, raw)
    for l in range(3):
        rows, cols = rows0 >> l, cols0 >> l
        assert depth[l].shape == (rows, cols)
        s_ = float(1 << l)
        f, cx, cy = c["f"] / s_, c["cx"] / s_, c["cy"] / s_
        want = {
            3: depth[l],
            4: gray[l],
        }
        dx, dy = O.derivative_image(gray[l])
        want[5], want[6] = dx, dy
        if l > 0:  # level 0 of the frame now holds its raycast model
            v = O.compute_vmap(depth[l], f, f, cx, cy, 3.0)
            want[0], want[1] = v, O.compute_nmap(v)
        for which, ref in want.items():
            ch = 4 if which in (0, 1) else 1
            got = download(g.frame_image(which, l), rows, cols, ref.dtype, ch)
            assert bits_equal(got, ref), (l, which, mismatch_report(got, ref))
    # The previous frame: level 0 = its raycast model, levels 1-2 = ResizeMap.
    v0 = download(g.frame_image(16 + 0, 0), rows0, cols0, np.float32, 4)
    n0 = download(g.frame_image(16 + 1, 0), rows0, cols0, np.float32, 4)
    v, n = v0, n0
    for l in (1, 2):
        v, n = O.resize_map(v, n)
        rows, cols = rows0 >> l, cols0 >> l
        gv = download(g.frame_image(16 + 0, l), rows, cols, np.float32, 4)
        gn = download(g.frame_image(16 + 1, l), rows, cols, np.float32, 4)
        assert bits_equal(gv, v), (l, mismatch_report(gv, v))
        assert bits_equal(gn, n), (l, mismatch_report(gn, n))
    assert np.isfinite(v0[..., 2]).sum() > 1000  # the model is not empty


def test_system_lost_frame_keeps_model(dev):
    """An all-zero depth frame cannot be tracked (count below MIN_ICP_COUNT):
    the pose rolls back, nothing is fused, and the next good frame tracks
    against the previous model (Tracking.cc:452-456)."""
    c = CFG_SMALL
    cfg = sys_cfg(c)
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    s = synth(c)
    seq = [0, 1, None, 2]
    for i in seq:
        if i is None:
            d = np.zeros((c["rows"], c["cols"]), np.uint16)
            rgb = np.zeros((c["rows"], c["cols"], 3), np.uint8)
            lab = np.zeros((c["rows"], c["cols"]), np.uint8)
        else:
            d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        rg, ro = g.result(), o.grab(d, rgb, lab)
        assert rg["tracked"] == ro["tracked"]
        if i is None:
            assert rg["tracked"] == 0
        assert np.abs(rg["pose"] - ro["pose"]).max() < 1e-4


@pytest.mark.parametrize("c", [CFG_SMALL, CFG_VGA])
def test_graph_form_matches_eager(dev, c):
    """GrabImage as one hipGraph launch per frame (staged inputs, two graphs
    by frame parity) gives the same result as the eager launches, frame by
    frame, with eager work and synchronisations between the replays; the
    final maps are identical."""
    import torch
    from helpers import bits_equal, download, mismatch_report
    cfg = sys_cfg(c)
    ge, gg = B.System(cfg), B.System(cfg)
    s = synth(c)
    frames = 9 if c is CFG_SMALL else 5
    for i in range(frames):
        d, rgb, lab = s.render(i)
        t = [to_dev(a, dev) for a in (d, rgb, lab)]
        ge.grab(*t)
        gg.grab_staged(*t, use_graph=True)
        x = torch.randn(1 << 16, device=dev)  # unrelated eager work between replays
        (x * 2).sum().item()
        re, rg = ge.result(), gg.result()
        assert re["tracked"] == rg["tracked"], i
        assert np.array_equal(re["pose"], rg["pose"]), (i, np.abs(re["pose"] - rg["pose"]).max())
        assert re["allocated_blocks"] == rg["allocated_blocks"], i
        assert re["visible_blocks"] == rg["visible_blocks"], i
    rows, cols = c["rows"], c["cols"]
    for which in (0, 1):
        a = download(ge.frame_image(which, 0), rows, cols, np.float32, 4)
        b = download(gg.frame_image(which, 0), rows, cols, np.float32, 4)
        assert bits_equal(a, b), (which, mismatch_report(a, b))
    ma, mb = B.map_stats(ge.map_handle()), B.map_stats(gg.map_handle())
    for k in ("allocated_blocks", "updated_total", "visible_total", "fusions"):
        assert ma[k] == mb[k], k


def test_system_rejects_bad_flags(dev):
    """The 0/1 switches of the configuration refuse other values."""
    for field in ("icp_only", "icp_sums_fp32", "icp_spread"):
        cfg = sys_cfg(CFG_SMALL)
        setattr(cfg, field, 2)
        h = B.C.c_void_p()
        assert B.lib().sicp_system_create(B.C.byref(cfg), B.C.byref(h)) != 0, field
        assert field in B.lib().sicp_last_error().decode(), field
