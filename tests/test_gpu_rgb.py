"""Parity of RGBStep (Tracking/Reduction.cu:387-642), SO3Step (:67-211) and
of tracking with the photometric term (ComputeSE3(false),
Tracking.cc:369-517) against the oracle.

The correspondence pass is integer work: its count and its int sum of
squared differences must match exactly, and so must sigma (formed from them
in float). The 29 row sums are FP32 products accumulated in FP64 in
different orders on the two sides, then rounded to float: the count must
match exactly and every other sum within 1e-6 of the largest sum of its kind
(|A|, |b|, r^2), as for ICP. Whole-system runs agree per frame to 1e-4 in
the pose and exactly in the tracked flag.
"""
import numpy as np
import pytest
import torch

from helpers import B, CFG_SMALL, CFG_VGA, K, sys_cfg, synth, to_dev
from oracle import binding as O

pytestmark = pytest.mark.gpu

REL = 1e-6


def _frame(c, i, level=0):
    d, rgb, _ = synth(c).render(i)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    v = O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    n = O.compute_nmap(v)
    g = O.image_to_intensity(rgb)
    for _ in range(level):
        v, n = O.resize_map(v, n)
        g = O.pyrdown_u8(g)
    dx, dy = O.derivative_image(g)
    return v, g, dx, dy


def _gpu_rgb(dev, ni, li, nv, lv, dx, dy, pc, pl, k, pad=0):
    t = [to_dev(a, dev, pad_cols=pad) for a in (ni, li, nv, lv, dx, dy)]
    rows, cols = ni.shape
    nbytes = B.lib().sicp_rgb_scratch_size(cols, rows)
    scratch = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    sums = torch.zeros(29, dtype=torch.float32, device=dev)
    hs = np.zeros(29, np.float32)
    corr = np.zeros(2, np.int32)
    sig = np.zeros(1, np.float32)
    A = np.zeros(36)
    b = np.zeros(6)
    res = np.zeros(2, np.float32)
    p = lambda a: a.ctypes.data_as(B.C.c_void_p)
    B.check(B.lib().sicp_rgb_step(*[B.image(x) for x in t], B.C.byref(pc), B.C.byref(pl), k,
                                  scratch.data_ptr(), nbytes, sums.data_ptr(), p(hs), p(corr),
                                  p(sig), p(A), p(b), p(res), B.stream_ptr()))
    return hs, (int(corr[0]), int(corr[1])), float(sig[0]), A.reshape(6, 6), b, res


def _check_sums(g, o):
    assert g[28] == o[28], (g[28], o[28])
    idxA = [i for i in range(27) if i not in (6, 12, 17, 21, 24, 26)]
    idxb = [6, 12, 17, 21, 24, 26]
    for idx in (idxA, idxb, [27]):
        scale = np.abs(o[idx]).max() + 1e-30
        err = np.abs(g[idx].astype(np.float64) - o[idx]).max() / scale
        assert err < REL, (idx, err)


@pytest.mark.parametrize("c", [CFG_SMALL, CFG_VGA])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_rgb_step_vs_oracle(dev, c, level):
    v0, g0, _, _ = _frame(c, 0, level)
    v1, g1, dx1, dy1 = _frame(c, 3, level)
    pc = B.pose_f(synth(c).pose(3))
    pl = B.pose_f(np.eye(4))
    hs, corr, sig, A, b, res = _gpu_rgb(dev, g1, g0, v1, v0, dx1, dy1, pc, pl, K(c, level))
    os_, ocorr, osig = O.rgb_sums(g1, g0, v1, v0, dx1, dy1, pc, pl, K(c, level))
    assert corr == ocorr and corr[0] > 300
    assert np.float32(sig).tobytes() == np.float32(osig).tobytes()
    _check_sums(hs, os_)
    assert np.array_equal(A, A.T)
    assert res[1] == hs[28] and res[0] == hs[27]
    Ao, bo = O.create_matrix(hs)
    assert np.array_equal(A, Ao) and np.array_equal(b, bo)


def test_rgb_step_pitched_inputs(dev):
    c = CFG_SMALL
    v0, g0, _, _ = _frame(c, 0)
    v1, g1, dx1, dy1 = _frame(c, 2)
    pc = B.pose_f(synth(c).pose(2))
    pl = B.pose_f(np.eye(4))
    a = _gpu_rgb(dev, g1, g0, v1, v0, dx1, dy1, pc, pl, K(c))
    p = _gpu_rgb(dev, g1, g0, v1, v0, dx1, dy1, pc, pl, K(c), pad=24)
    assert np.array_equal(a[0], p[0]) and a[1] == p[1] and a[2] == p[2]


def test_rgb_step_identical_frames(dev):
    """Zero motion on identical frames: every difference is 0, so sigma is 0,
    the weight 1, b = 0 and sum r^2 = 0."""
    c = CFG_SMALL
    v, g, dx, dy = _frame(c, 0)
    p = B.pose_f(np.eye(4))
    hs, corr, sig, A, b, res = _gpu_rgb(dev, g, g, v, v, dx, dy, p, p, K(c))
    _, ocorr, _ = O.rgb_sums(g, g, v, v, dx, dy, p, p, K(c))
    assert corr == ocorr and corr[1] == 0 and sig == 0.0
    assert hs[28] == corr[0] and hs[27] == 0.0 and np.all(b == 0.0)
    w = np.linalg.eigvalsh(A)
    assert w.min() > -1e-6 * w.max()


def test_rgb_step_rejects_bad_arguments(dev):
    c = CFG_SMALL
    v, g, dx, dy = _frame(c, 0)
    p = B.pose_f(np.eye(4))
    t = [to_dev(a, dev) for a in (g, g, v, v, dx, dy)]
    small = torch.zeros(1024, dtype=torch.uint8, device=dev)
    sums = torch.zeros(29, dtype=torch.float32, device=dev)
    rc = B.lib().sicp_rgb_step(*[B.image(x) for x in t], B.C.byref(p), B.C.byref(p), K(c),
                               small.data_ptr(), 1024, sums.data_ptr(), None, None, None, None,
                               None, None, B.stream_ptr())
    assert rc != 0
    t2 = list(t)
    t2[4] = to_dev(dx[:-1], dev)
    rc = B.lib().sicp_rgb_step(*[B.image(x) for x in t2], B.C.byref(p), B.C.byref(p), K(c),
                               small.data_ptr(), 1024, sums.data_ptr(), None, None, None, None,
                               None, None, B.stream_ptr())
    assert rc != 0


def _gpu_so3(dev, ni, li, dx, dy, pc, pl, k, pad=0):
    t = [to_dev(a, dev, pad_cols=pad) for a in (ni, li, dx, dy)]
    partials = torch.zeros(B.lib().sicp_icp_partials_size(), dtype=torch.uint8, device=dev)
    sums = torch.zeros(11, dtype=torch.float32, device=dev)
    hs = np.zeros(11, np.float32)
    A = np.zeros(9)
    b = np.zeros(3)
    res = np.zeros(2, np.float32)
    p = lambda a: a.ctypes.data_as(B.C.c_void_p)
    B.check(B.lib().sicp_so3_step(*[B.image(x) for x in t], B.C.byref(pc), B.C.byref(pl), k,
                                  partials.data_ptr(), sums.data_ptr(), p(hs), p(A), p(b),
                                  p(res), B.stream_ptr()))
    return hs, A.reshape(3, 3), b, res


@pytest.mark.parametrize("c", [CFG_SMALL, CFG_VGA])
@pytest.mark.parametrize("level", [0, 2])
def test_so3_step_vs_oracle(dev, c, level):
    """ComputeSO3's call (level 2) and level 0: count exact, the other sums
    within 1e-6 of the largest of their kind; pitched inputs give the same
    sums."""
    _, g0, _, _ = _frame(c, 0, level)
    _, g1, dx1, dy1 = _frame(c, 3, level)
    pc = B.pose_f(synth(c).pose(3))
    pl = B.pose_f(np.eye(4))
    hs, A, b, res = _gpu_so3(dev, g1, g0, dx1, dy1, pc, pl, K(c, level))
    os_ = O.so3_sums(g1, g0, dx1, dy1, pc, pl, K(c, level))
    assert hs[10] == os_[10] and os_[10] > 500
    idxA, idxb = [0, 1, 2, 4, 5, 7], [3, 6, 8]
    for idx in (idxA, idxb, [9]):
        scale = np.abs(os_[idx]).max() + 1e-30
        assert np.abs(hs[idx].astype(np.float64) - os_[idx]).max() / scale < REL, idx
    Ao, bo = O.create_matrix3(hs)
    assert np.array_equal(A, Ao) and np.array_equal(b, bo)
    assert res[0] == hs[9] and res[1] == hs[10]
    hp = _gpu_so3(dev, g1, g0, dx1, dy1, pc, pl, K(c, level), pad=40)[0]
    assert np.array_equal(hs, hp)


def _rgb_cfg(c, k32=0):
    cfg = sys_cfg(c)
    cfg.icp_only = 0
    cfg.icp_sums_fp32 = k32
    return cfg


# k32 = 1: the precision form's FP32 compensated sums (photometric and ICP)
# on both sides.
@pytest.mark.parametrize("c,frames,k32", [(CFG_SMALL, 12, 0), (CFG_VGA, 4, 0), (CFG_SMALL, 8, 1)])
def test_system_rgb_vs_oracle(dev, c, frames, k32):
    cfg = _rgb_cfg(c, k32)
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    s = synth(c)
    for i in range(frames):
        d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        rg = g.result()
        ro = o.grab(d, rgb, lab)
        assert rg["tracked"] == ro["tracked"], (i, rg, ro)
        assert rg["iterations"] == ro["iterations"], (i, rg["iterations"], ro["iterations"])
        dp = np.abs(rg["pose"] - ro["pose"]).max()
        assert dp < 1e-4, (i, dp)
        assert abs(rg["rgb_count"] - ro["rgb_count"]) <= 0.001 * ro["rgb_count"] + 2, i
        if i > 0:
            assert ro["rgb_count"] > 1000
        na, nb = rg["allocated_blocks"], ro["allocated_blocks"]
        assert abs(na - nb) <= 0.01 * nb + 2, (i, na, nb)


def test_system_rgb_lost_frame_keeps_last_images(dev):
    """A black, depthless frame fails the photometric test first ("track rgb
    failed", Tracking.cc:427-431): the pose rolls back and the next frame is
    tracked against the previous frame's model and intensities."""
    c = CFG_SMALL
    cfg = _rgb_cfg(c)
    g = B.System(cfg)
    o = O.OracleSystem(cfg)
    s = synth(c)
    for i in [0, 1, None, 2, 3]:
        if i is None:
            d = np.zeros((c["rows"], c["cols"]), np.uint16)
            rgb = np.zeros((c["rows"], c["cols"], 3), np.uint8)
            lab = np.zeros((c["rows"], c["cols"]), np.uint8)
        else:
            d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        rg, ro = g.result(), o.grab(d, rgb, lab)
        assert rg["tracked"] == ro["tracked"], i
        if i is None:
            assert rg["tracked"] == 0 and rg["rgb_count"] == 0
        else:
            assert rg["tracked"] == 1
        assert np.abs(rg["pose"] - ro["pose"]).max() < 1e-4, i


def test_system_rgb_graph_form_matches_eager(dev):
    c = CFG_SMALL
    cfg = _rgb_cfg(c)
    ge, gg = B.System(cfg), B.System(cfg)
    s = synth(c)
    for i in range(7):
        d, rgb, lab = s.render(i)
        t = [to_dev(a, dev) for a in (d, rgb, lab)]
        ge.grab(*t)
        gg.grab_staged(*t, use_graph=True)
        x = torch.randn(1 << 16, device=dev)
        (x * 2).sum().item()
        re, rg = ge.result(), gg.result()
        assert re["tracked"] == rg["tracked"], i
        assert np.array_equal(re["pose"], rg["pose"]), i
        assert re["rgb_count"] == rg["rgb_count"], i
