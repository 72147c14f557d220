"""Parity of the ICP reduction (Tracking/Reduction.cu:213-385) and of the
device Gauss-Newton loop against the oracle.

Both sides compute the per-pixel 29-vectors in FP32 in the reference's order
and accumulate them in FP64, in different orders, then round the totals to
float. The count must match exactly; every other sum within 1e-6 of the
largest sum of its kind (|A|, |b|, r^2) — a few float ulps.
"""
import numpy as np
import pytest
import torch

from helpers import B, CFG_SMALL, CFG_VGA, K, host, synth, to_dev
from oracle import binding as O

pytestmark = pytest.mark.gpu

REL = 1e-6


def _maps(c, frame, seed=0):
    d, _, _ = synth(c, seed=seed).render(frame)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    v = O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    return v, O.compute_nmap(v)


def _gpu_sums(dev, vc, vl, nc, nl, pc, pl, k, fp32=False):
    t = [to_dev(a, dev) for a in (vc, vl, nc, nl)]
    partials = torch.zeros(B.lib().sicp_icp_partials_size(), dtype=torch.uint8, device=dev)
    sums = torch.zeros(29, dtype=torch.float32, device=dev)
    hs = np.zeros(29, np.float32)
    A = np.zeros(36)
    b = np.zeros(6)
    res = np.zeros(2, np.float32)
    p = lambda a: a.ctypes.data_as(B.C.c_void_p)
    fn = B.lib().sicp_icp_step_fp32 if fp32 else B.lib().sicp_icp_step
    B.check(fn(B.image(t[0]), B.image(t[1]), B.image(t[2]), B.image(t[3]), B.C.byref(pc),
               B.C.byref(pl), k, partials.data_ptr(), sums.data_ptr(), p(hs), p(A), p(b), p(res),
               B.stream_ptr()))
    return hs, A.reshape(6, 6), b, res, sums


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
def test_icp_sums_vs_oracle(dev, c, level):
    v0, n0 = _maps(c, 0)
    v1, n1 = _maps(c, 3)
    for _ in range(level):
        v0, n0 = O.resize_map(v0, n0)
        v1, n1 = O.resize_map(v1, n1)
    s = synth(c)
    pc = B.pose_f(s.pose(3))
    pl = B.pose_f(np.eye(4))
    hs, A, b, res, _ = _gpu_sums(dev, v1, v0, n1, n0, pc, pl, K(c, level))
    os_ = O.icp_sums(v1, v0, n1, n0, pc, pl, K(c, level))
    _check_sums(hs, os_)
    assert os_[28] > 500
    # CreateMatrix layout: A mirrored, b = row sums, residual = (r^2, count)
    assert np.array_equal(A, A.T)
    assert res[1] == hs[28] and res[0] == hs[27]
    Ao, bo = O.create_matrix(hs)
    assert np.array_equal(A, Ao) and np.array_equal(b, bo)


@pytest.mark.parametrize("c", [CFG_SMALL, CFG_VGA])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_icp_sums_fp32_compensated_bit_exact(dev, c, level):
    """sicp_icp_step_fp32 (the precision form's FP32 compensated sums) against
    the oracle's restatement in the same order: bit-exact. Against the exact
    sums of the same FP32 products (math.fsum): within half an FP32 ulp of the
    sum plus 2^-40 of the sum of magnitudes, for every one of the 29."""
    import math
    v0, n0 = _maps(c, 0)
    v1, n1 = _maps(c, 3)
    for _ in range(level):
        v0, n0 = O.resize_map(v0, n0)
        v1, n1 = O.resize_map(v1, n1)
    pc = B.pose_f(synth(c).pose(3))
    pl = B.pose_f(np.eye(4))
    hs, A, b, res, _ = _gpu_sums(dev, v1, v0, n1, n0, pc, pl, K(c, level), fp32=True)
    _, rows = O.icp_sums(v1, v0, n1, n0, pc, pl, K(c, level), rows_out=True)
    ok = O.sums_k32(rows)
    assert np.array_equal(hs.view(np.uint32), ok.view(np.uint32)), np.abs(hs - ok).max()
    assert rows[:, 28].sum() > 500
    r64 = rows.astype(np.float64)
    for j in range(29):
        exact = math.fsum(r64[:, j])
        mag = math.fsum(np.abs(r64[:, j]))
        assert abs(float(hs[j]) - exact) <= 2.0 ** -24 * abs(exact) + 2.0 ** -40 * mag, j
    # the plain FP64 step agrees to the test's tolerance
    hd = _gpu_sums(dev, v1, v0, n1, n0, pc, pl, K(c, level))[0]
    _check_sums(hs, hd.astype(np.float64))


def test_icp_step_refuses_misaligned_partials(dev):
    """The partial rows are folded in 16-byte pairs: a scratch pointer off a
    16-byte boundary is refused before any launch."""
    c = CFG_SMALL
    v, n = _maps(c, 0)
    t = [to_dev(a, dev) for a in (v, v, n, n)]
    partials = torch.zeros(B.lib().sicp_icp_partials_size() + 16, dtype=torch.uint8, device=dev)
    sums = torch.zeros(29, dtype=torch.float32, device=dev)
    pc = B.pose_f(np.eye(4))
    for fn in (B.lib().sicp_icp_step, B.lib().sicp_icp_step_fp32):
        rc = fn(B.image(t[0]), B.image(t[1]), B.image(t[2]), B.image(t[3]), B.C.byref(pc),
                B.C.byref(pc), K(c), partials.data_ptr() + 8, sums.data_ptr(), None, None, None,
                None, B.stream_ptr())
        assert rc != 0
        assert "16-byte aligned" in B.lib().sicp_last_error().decode()


def test_icp_identical_frames(dev):
    """Zero motion on identical maps: r = 0 for every match, so b = 0, r^2 = 0
    and every valid vertex with a valid normal is a correspondence."""
    c = CFG_SMALL
    v, n = _maps(c, 0)
    pc = B.pose_f(np.eye(4))
    hs, A, b, res, _ = _gpu_sums(dev, v, v, n, n, pc, pc, K(c))
    valid = (~np.isnan(v[..., 0])) & (v[..., 2] >= 1e-3) & (~np.isnan(n[..., 0]))
    assert hs[28] == valid.sum()
    assert hs[27] == 0.0
    assert np.all(b == 0.0)
    w = np.linalg.eigvalsh(A)
    assert w.min() > -1e-6 * w.max()


def test_icp_deterministic(dev):
    c = CFG_VGA
    v0, n0 = _maps(c, 0)
    v1, n1 = _maps(c, 2)
    pc = B.pose_f(synth(c).pose(2))
    pl = B.pose_f(np.eye(4))
    for fp32 in (False, True):
        a = _gpu_sums(dev, v1, v0, n1, n0, pc, pl, K(c), fp32)[0]
        b = _gpu_sums(dev, v1, v0, n1, n0, pc, pl, K(c), fp32)[0]
        assert np.array_equal(a, b)


def test_se3_update_host_matches_oracle():
    rng = np.random.default_rng(3)
    M = rng.normal(size=(6, 6))
    A = M @ M.T + 6 * np.eye(6)
    b = rng.normal(size=6) * 1e-2
    last = O.se3_exp(rng.normal(size=6) * 0.1)
    delta = np.eye(4)
    nxt = np.zeros(16)
    x = np.zeros(6)
    p = lambda a: a.ctypes.data_as(B.C.c_void_p)
    d = delta.copy()
    B.check(B.lib().sicp_se3_update_host(p(A), p(b), p(d), p(last), p(nxt), p(x)))
    od, on, ox = O.se3_update(A, b, delta, last)
    assert np.allclose(x, ox, rtol=0, atol=1e-15)
    assert np.allclose(nxt.reshape(4, 4), on, rtol=0, atol=1e-14)
