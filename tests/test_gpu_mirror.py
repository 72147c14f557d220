"""The reference-named C++ layer (semantic_icpslam_amd/include/sicp_ref:
DeviceArray2D, Intrinsics, Matrix3f, DeviceMap, FilterDepth ... ICPStep,
FuseMapColor, CheckBlockVisibility, CreateRenderingBlocks, Raycast), driven
the way the reference's Frame / Tracker / Mapping call it, against the
oracle on the same seeded frames. Pyramid, map and raycast results are
bit-exact; the ICP system is within the ICP reduction's tolerance (1e-6 of
the largest entry of its kind)."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import B, CFG_SMALL, K, bits_equal, map_cfg, mismatch_report, synth
from oracle import binding as O

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libmirror_check.so")


def _lib():
    assert os.path.exists(LIB), "tests/cpp/libmirror_check.so not built (run build())"
    return C.CDLL(LIB)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _arr3(arrs):
    return (C.c_void_p * 3)(*[_p(a) for a in arrs])


def _pose12(T):
    T = np.asarray(T, np.float64)
    return np.concatenate([T[:3, :3].astype(np.float32).ravel(),
                           T[:3, 3].astype(np.float32)]).astype(np.float32)


@pytest.mark.parametrize("c", [CFG_SMALL, dict(CFG_SMALL, cols=317, rows=239, cx=158.0, cy=119.0)])
def test_mirror_frame_pyramid(dev, c):
    d, rgb, _ = synth(c).render(1)
    cols, rows = c["cols"], c["rows"]
    dims = [(rows >> l, cols >> l) for l in range(3)]
    raw = np.zeros((rows, cols), np.float32)
    depth = [np.zeros(s, np.float32) for s in dims]
    gray = [np.zeros(s, np.uint8) for s in dims]
    dx = [np.zeros(s, np.int16) for s in dims]
    dy = [np.zeros(s, np.int16) for s in dims]
    vm = [np.zeros(s + (4,), np.float32) for s in dims]
    nm = [np.zeros(s + (4,), np.float32) for s in dims]
    vr = [np.zeros(s + (4,), np.float32) for s in dims]
    nr = [np.zeros(s + (4,), np.float32) for s in dims]
    f = _lib().mirror_frame
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int] + [C.c_float] * 6 + [C.c_void_p] + \
        [C.c_void_p] * 8
    rc = f(_p(np.ascontiguousarray(d)), _p(np.ascontiguousarray(rgb)), cols, rows, c["f"], c["f"],
           c["cx"], c["cy"], 1000.0, 3.0, _p(raw), _arr3(depth), _arr3(gray), _arr3(dx),
           _arr3(dy), _arr3(vm), _arr3(nm), _arr3(vr), _arr3(nr))
    assert rc == 0
    oraw, ofilt = O.filter_depth(d, 1000.0, 3.0)
    assert bits_equal(raw, oraw), mismatch_report(raw, oraw)
    od, og = [ofilt], [O.image_to_intensity(rgb)]
    for _ in range(2):
        od.append(O.pyrdown_f32(od[-1]))
        og.append(O.pyrdown_u8(og[-1]))
    ov, on = [], []
    for l in range(3):
        s_ = float(1 << l)
        v = O.compute_vmap(od[l], c["f"] / s_, c["f"] / s_, c["cx"] / s_, c["cy"] / s_, 3.0)
        ov.append(v)
        on.append(O.compute_nmap(v))
        odx, ody = O.derivative_image(og[l])
        assert bits_equal(depth[l], od[l]), (l, mismatch_report(depth[l], od[l]))
        assert np.array_equal(gray[l], og[l]), l
        assert np.array_equal(dx[l], odx) and np.array_equal(dy[l], ody), l
        assert bits_equal(vm[l], ov[l]), (l, mismatch_report(vm[l], ov[l]))
        assert bits_equal(nm[l], on[l]), (l, mismatch_report(nm[l], on[l]))
    v, n = ov[0], on[0]
    for l in (1, 2):
        v, n = O.resize_map(v, n)
        assert bits_equal(vr[l], v) and bits_equal(nr[l], n), l


def test_mirror_icp_step(dev):
    c = CFG_SMALL
    s = synth(c)
    maps = []
    for i in (0, 3):
        d, _, _ = s.render(i)
        _, filt = O.filter_depth(d, 1000.0, 3.0)
        v = O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0)
        maps.append((v, O.compute_nmap(v)))
    (v0, n0), (v1, n1) = maps
    Tc, Tl = s.pose(3), np.eye(4)
    A = np.zeros(36)
    b = np.zeros(6)
    res = np.zeros(2, np.float32)
    f = _lib().mirror_icp
    f.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_float] * 4 + [C.c_void_p] * 5
    rc = f(_p(v1), _p(v0), _p(n1), _p(n0), c["cols"], c["rows"], c["f"], c["f"], c["cx"], c["cy"],
           _p(_pose12(Tc)), _p(_pose12(Tl)), _p(A), _p(b), _p(res))
    assert rc == 0
    os_ = O.icp_sums(v1, v0, n1, n0, B.pose_f(Tc), B.pose_f(Tl), K(c))
    Ao, bo = O.create_matrix(os_)
    assert res[1] == os_[28] and os_[28] > 500
    assert np.abs(A.reshape(6, 6) - Ao).max() <= 1e-6 * np.abs(Ao).max()
    assert np.abs(b - bo).max() <= 1e-6 * np.abs(bo).max()
    assert abs(res[0] - os_[27]) <= 1e-6 * os_[27]


def test_mirror_rgb_step(dev):
    c = CFG_SMALL
    s = synth(c)
    fr = []
    for i in (0, 3):
        d, rgb, _ = s.render(i)
        _, filt = O.filter_depth(d, 1000.0, 3.0)
        v = O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0)
        g = O.image_to_intensity(rgb)
        fr.append((v, g) + tuple(O.derivative_image(g)))
    (v0, g0, _, _), (v1, g1, dx1, dy1) = fr
    Tc, Tl = s.pose(3), np.eye(4)
    A = np.zeros(36)
    b = np.zeros(6)
    res = np.zeros(2, np.float32)
    corr = np.zeros(2, np.int32)
    f = _lib().mirror_rgb
    f.argtypes = [C.c_void_p] * 6 + [C.c_int, C.c_int] + [C.c_float] * 4 + [C.c_void_p] * 6
    rc = f(_p(g1), _p(g0), _p(v1), _p(v0), _p(dx1), _p(dy1), c["cols"], c["rows"], c["f"],
           c["f"], c["cx"], c["cy"], _p(_pose12(Tc)), _p(_pose12(Tl)), _p(A), _p(b), _p(res),
           _p(corr))
    assert rc == 0
    os_, ocorr, _ = O.rgb_sums(g1, g0, v1, v0, dx1, dy1, B.pose_f(Tc), B.pose_f(Tl), K(c))
    Ao, bo = O.create_matrix(os_)
    assert tuple(corr) == ocorr and res[1] == os_[28] and os_[28] > 500
    assert np.abs(A.reshape(6, 6) - Ao).max() <= 1e-6 * np.abs(Ao).max()
    assert np.abs(b - bo).max() <= 1e-6 * np.abs(bo).max()
    assert abs(res[0] - os_[27]) <= 1e-6 * os_[27]


def test_mirror_so3_step(dev):
    c = CFG_SMALL
    s = synth(c)
    gs = []
    for i in (0, 3):
        _, rgb, _ = s.render(i)
        gs.append(O.image_to_intensity(rgb))
    g0, g1 = gs
    dx1, dy1 = O.derivative_image(g1)
    Tc, Tl = s.pose(3), np.eye(4)
    A = np.zeros(9)
    b = np.zeros(3)
    res = np.zeros(2, np.float32)
    f = _lib().mirror_so3
    f.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_float] * 4 + [C.c_void_p] * 5
    rc = f(_p(g1), _p(g0), _p(dx1), _p(dy1), c["cols"], c["rows"], c["f"], c["f"], c["cx"],
           c["cy"], _p(_pose12(Tc)), _p(_pose12(Tl)), _p(A), _p(b), _p(res))
    assert rc == 0
    os_ = O.so3_sums(g1, g0, dx1, dy1, B.pose_f(Tc), B.pose_f(Tl), K(c))
    Ao, bo = O.create_matrix3(os_)
    assert res[1] == os_[10] and os_[10] > 500
    assert np.abs(A.reshape(3, 3) - Ao).max() <= 1e-6 * np.abs(Ao).max()
    assert np.abs(b - bo).max() <= 1e-6 * np.abs(bo).max()


def test_mirror_fuse_and_raycast(dev):
    c = CFG_SMALL
    cfg = map_cfg(c)
    s = synth(c)
    frames = [0, 2, 4]
    om = O.OracleMap(cfg)
    ins = []
    ovis = []
    for i in frames:
        d, rgb, lab = s.render(i)
        raw, filt = O.filter_depth(d, 1000.0, 3.0)
        n = O.compute_nmap(O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0))
        pose = B.pose_f(s.pose(i))
        ovis.append(om.fuse(raw, rgb, n, lab, 1, pose, K(c)))
        ins.append((raw, rgb, n, lab, _pose12(s.pose(i))))
    cols, rows = c["cols"], c["rows"]
    stack = lambda k: np.ascontiguousarray(np.stack([x[k] for x in ins]))
    zc, zr = (cols + 7) // 8, (rows + 7) // 8
    vis = np.zeros(len(frames), np.uint32)
    chk = C.c_uint32(0)
    rendered = C.c_int(0)
    zmin = np.zeros((zr, zc), np.float32)
    zmax = np.zeros((zr, zc), np.float32)
    vm = np.zeros((rows, cols, 4), np.float32)
    nm = np.zeros((rows, cols, 4), np.float32)
    max_tris = 400000
    mesh_v = np.zeros((max_tris, 3, 3), np.float32)
    mesh_c = np.zeros((max_tris, 3, 3), np.uint8)
    ntris = C.c_uint32(0)
    f = _lib().mirror_map
    f.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5 + [C.c_int, C.c_int] + \
        [C.c_float] * 4 + [C.c_void_p] * 7 + [C.c_uint32] + [C.c_void_p] * 3
    rc = f(C.byref(cfg), len(frames), _p(stack(0)), _p(stack(1)), _p(stack(2)), _p(stack(3)),
           _p(stack(4)), cols, rows, c["f"], c["f"], c["cx"], c["cy"], _p(vis), C.byref(chk),
           C.byref(rendered), _p(zmin), _p(zmax), _p(vm), _p(nm), max_tris, _p(mesh_v),
           _p(mesh_c), C.byref(ntris))
    assert rc == 0
    ov_, _, oc_, ototal = om.mesh(B.mesh_tables(), max_tris)
    assert ntris.value == ototal == len(ov_) and ototal > 1000
    assert bits_equal(mesh_v[: ntris.value], ov_)
    assert np.array_equal(mesh_c[: ntris.value], oc_)
    assert list(vis) == ovis
    pose = B.pose_f(s.pose(frames[-1]))
    assert chk.value == om.visibility(pose, cols, rows, K(c), 3.0, 0.1)
    ozmin, ozmax, oany = om.rendering_blocks(zc, zr, 3.0, 0.1, pose, K(c))
    assert rendered.value == 1 and oany
    assert bits_equal(zmin, ozmin) and bits_equal(zmax, ozmax)
    ov, on = om.raycast(cols, rows, ozmin, ozmax, pose, K(c))
    assert bits_equal(vm, ov), mismatch_report(vm, ov)
    assert bits_equal(nm, on), mismatch_report(nm, on)
    assert (~np.isnan(ov[..., 0])).mean() > 0.3
