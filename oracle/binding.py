"""ctypes binding of the oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
bench.py, as the checker or the timed CPU baseline; never by the product.
All arrays are contiguous numpy host arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from semantic_icpslam_amd.binding import (  # noqa: E402
    BLOCK, HASH_ENTRY, VOXEL_GEOM, VOXEL_SEM, FrameResult, Intrinsics, MapConfig, MapStats,
    PoseF, SystemConfig)

# ORACLE_LIB names another build of the same sources (the sanitizer builds).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None
_P = C.c_void_p

_SIGS = {
    "oracle_set_threads": (C.c_int, [C.c_int]),
    "oracle_max_threads": (C.c_int, []),
    "oracle_expf": (C.c_float, [C.c_float]),
    "oracle_filter_depth": (None, [_P, C.c_int, C.c_int, C.c_float, C.c_float, _P, _P]),
    "oracle_compute_vmap": (None, [_P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float, _P]),
    "oracle_compute_nmap": (None, [_P, C.c_int, C.c_int, _P]),
    "oracle_pyrdown_f32": (None, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int]),
    "oracle_pyrdown_u8": (None, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int]),
    "oracle_image_to_intensity": (None, [_P, C.c_int, C.c_int, _P]),
    "oracle_derivative_image": (None, [_P, C.c_int, C.c_int, _P, _P]),
    "oracle_resize_map": (None, [_P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int]),
    "oracle_icp_sums": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.POINTER(PoseF),
                               C.POINTER(PoseF), Intrinsics, _P, _P]),
    "oracle_icp_sums_fp64": (None, [_P, _P, _P, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
    "oracle_sums_k32": (None, [_P, C.c_int, C.c_int, _P]),
    "oracle_rgb_sums": (None, [_P, _P, _P, _P, _P, _P, C.c_int, C.c_int, C.POINTER(PoseF),
                               C.POINTER(PoseF), Intrinsics, _P, _P, _P]),
    "oracle_so3_sums": (None, [_P, _P, _P, _P, C.c_int, C.c_int, C.POINTER(PoseF),
                               C.POINTER(PoseF), Intrinsics, _P]),
    "oracle_create_matrix": (None, [_P, _P, _P]),
    "oracle_cholesky_solve6": (C.c_int, [_P, _P, _P]),
    "oracle_se3_exp": (None, [_P, _P]),
    "oracle_se3_update": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "oracle_euler_sin_norm": (None, [_P, _P, _P, _P]),
    "oracle_map_create": (_P, [C.POINTER(MapConfig), C.c_int]),
    "oracle_map_destroy": (None, [_P]),
    "oracle_map_reset": (None, [_P]),
    "oracle_map_hash": (C.c_uint32, [_P, C.c_int, C.c_int, C.c_int]),
    "oracle_fuse_map_color": (C.c_uint32, [_P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(PoseF), Intrinsics]),
    "oracle_fuse_map_color_fp64": (C.c_uint32, [_P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int,
                                                _P, _P]),
    "oracle_check_visibility": (C.c_uint32, [_P, C.POINTER(PoseF), C.c_int, C.c_int, Intrinsics,
                                             C.c_float, C.c_float]),
    "oracle_create_rendering_blocks": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_float,
                                                 C.c_float, C.POINTER(PoseF), Intrinsics]),
    "oracle_raycast": (None, [_P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_int, C.POINTER(PoseF),
                              Intrinsics]),
    "oracle_map_get_stats": (None, [_P, C.POINTER(MapStats)]),
    "oracle_map_export": (None, [_P, _P, _P, _P, _P, _P]),
    "oracle_map_visible": (C.c_uint32, [_P, _P, C.c_uint32]),
    "oracle_mesh_scene": (C.c_uint32, [_P, _P, _P, _P, C.c_uint32, _P, _P, _P, _P]),
    "oracle_unify_neighbor": (None, [_P]),
    "oracle_map_clone_fp64": (_P, [_P]),
    "oracle_f32_to_f16": (C.c_uint16, [C.c_float]),
    "oracle_f16_to_f32": (C.c_float, [C.c_uint16]),
    "oracle_raycast_fp64": (None, [_P, _P, _P, C.c_int, C.c_int, _P, _P, C.c_int, _P, C.c_double,
                                   C.c_double, C.c_double, C.c_double]),
    "oracle_map_get_sdf": (C.c_double, [_P, C.c_int, C.c_int]),
    "oracle_map_set_sdf": (None, [_P, C.c_int, C.c_int, C.c_double]),
    "oracle_map_set_label": (None, [_P] + [C.c_int] * 7),
    "oracle_map_counters": (None, [_P, _P]),
    "oracle_system_create": (_P, [C.POINTER(SystemConfig), C.c_int]),
    "oracle_system_destroy": (None, [_P]),
    "oracle_system_grab": (C.c_int, [_P, _P, _P, _P, C.c_int, C.POINTER(FrameResult)]),
    "oracle_system_model": (None, [_P, C.c_int, _P, _P]),
    "oracle_system_work": (None, [_P, _P]),
    "oracle_se3_control_trace": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_float, C.c_int, _P,
                                           C.POINTER(C.c_float)]),
    "oracle_se3_gate": (C.c_int, [C.c_float, C.c_float, _P, _P]),
    "oracle_track_frame": (C.c_int, [C.POINTER(SystemConfig), C.c_int, _P, _P, _P, _P, _P, _P,
                                     C.POINTER(FrameResult)]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_threads(n):
    return lib().oracle_set_threads(n)


def expf(x):
    return lib().oracle_expf(float(x))


def filter_depth(depth_u16, scale, cutoff):
    d = _c(depth_u16, np.uint16)
    rows, cols = d.shape
    raw = np.empty((rows, cols), np.float32)
    filt = np.empty((rows, cols), np.float32)
    lib().oracle_filter_depth(_p(d), cols, rows, scale, cutoff, _p(raw), _p(filt))
    return raw, filt


def compute_vmap(depth, fx, fy, cx, cy, cutoff):
    d = _c(depth, np.float32)
    rows, cols = d.shape
    v = np.empty((rows, cols, 4), np.float32)
    lib().oracle_compute_vmap(_p(d), cols, rows, fx, fy, cx, cy, cutoff, _p(v))
    return v


def compute_nmap(vmap):
    v = _c(vmap, np.float32)
    rows, cols = v.shape[:2]
    n = np.empty_like(v)
    lib().oracle_compute_nmap(_p(v), cols, rows, _p(n))
    return n


def pyrdown_f32(src):
    s = _c(src, np.float32)
    rows, cols = s.shape
    d = np.empty((rows // 2, cols // 2), np.float32)
    lib().oracle_pyrdown_f32(_p(s), cols, rows, _p(d), cols // 2, rows // 2)
    return d


def pyrdown_u8(src):
    s = _c(src, np.uint8)
    rows, cols = s.shape
    d = np.empty((rows // 2, cols // 2), np.uint8)
    lib().oracle_pyrdown_u8(_p(s), cols, rows, _p(d), cols // 2, rows // 2)
    return d


def image_to_intensity(rgb):
    s = _c(rgb, np.uint8)
    rows, cols = s.shape[:2]
    g = np.empty((rows, cols), np.uint8)
    lib().oracle_image_to_intensity(_p(s), cols, rows, _p(g))
    return g


def derivative_image(gray):
    s = _c(gray, np.uint8)
    rows, cols = s.shape
    dx = np.empty((rows, cols), np.int16)
    dy = np.empty((rows, cols), np.int16)
    lib().oracle_derivative_image(_p(s), cols, rows, _p(dx), _p(dy))
    return dx, dy


def resize_map(vsrc, nsrc):
    v = _c(vsrc, np.float32)
    n = _c(nsrc, np.float32)
    rows, cols = v.shape[:2]
    vd = np.empty((rows // 2, cols // 2, 4), np.float32)
    nd = np.empty_like(vd)
    lib().oracle_resize_map(_p(v), _p(n), cols, rows, _p(vd), _p(nd), cols // 2, rows // 2)
    return vd, nd


def icp_sums(vcurr, vlast, ncurr, nlast, pose_curr, pose_last, K, rows_out=False):
    a = [_c(x, np.float32) for x in (vcurr, vlast, ncurr, nlast)]
    rows, cols = a[0].shape[:2]
    sums = np.zeros(29, np.float64)
    ro = np.zeros((rows * cols, 29), np.float32) if rows_out else None
    lib().oracle_icp_sums(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), cols, rows,
                          C.byref(pose_curr), C.byref(pose_last), K, _p(sums), _p(ro))
    return (sums, ro) if rows_out else sums


def sums_k32(rows29, compensate=True):
    """The precision form's FP32 compensated sums of per-pixel rows (N x 29,
    slot 28 = 1 where the pixel contributes), in the GPU reduction's order;
    compensate=False: plain FP32 in the same order."""
    r = _c(rows29, np.float32)
    assert r.ndim == 2 and r.shape[1] == 29
    out = np.zeros(29, np.float32)
    lib().oracle_sums_k32(_p(r), r.shape[0], 1 if compensate else 0, _p(out))
    return out


def icp_sums_fp64(vcurr, vlast, ncurr, nlast, T_curr, T_last, K4):
    a = [_c(x, np.float64) for x in (vcurr, vlast, ncurr, nlast)]
    rows, cols = a[0].shape[:2]
    sums = np.zeros(29, np.float64)
    Tc = _c(T_curr, np.float64)
    Tl = _c(T_last, np.float64)
    k = _c(K4, np.float64)
    lib().oracle_icp_sums_fp64(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), cols, rows, _p(Tc),
                               _p(Tl), _p(k), _p(sums))
    return sums


def rgb_sums(nimg, limg, nvmap, lvmap, dx, dy, pose_curr, pose_last, K):
    """RGBStep: (29 FP64 sums, (count, sum of squared differences), sigma)."""
    ni, li = _c(nimg, np.uint8), _c(limg, np.uint8)
    nv, lv = _c(nvmap, np.float32), _c(lvmap, np.float32)
    gx, gy = _c(dx, np.int16), _c(dy, np.int16)
    rows, cols = ni.shape[:2]
    for a in (li, gx, gy):
        assert a.shape[:2] == (rows, cols)
    assert nv.shape == lv.shape == (rows, cols, 4)
    sums = np.zeros(29, np.float64)
    corr = np.zeros(2, np.int32)
    sig = np.zeros(1, np.float32)
    lib().oracle_rgb_sums(_p(ni), _p(li), _p(nv), _p(lv), _p(gx), _p(gy), cols, rows,
                          C.byref(pose_curr), C.byref(pose_last), K, _p(sums), _p(corr),
                          _p(sig))
    return sums, (int(corr[0]), int(corr[1])), float(sig[0])


def so3_sums(nimg, limg, dx, dy, pose_curr, pose_last, K):
    """SO3Step: the 11 FP64 sums."""
    ni, li = _c(nimg, np.uint8), _c(limg, np.uint8)
    gx, gy = _c(dx, np.int16), _c(dy, np.int16)
    rows, cols = ni.shape[:2]
    for a in (li, gx, gy):
        assert a.shape[:2] == (rows, cols)
    sums = np.zeros(11, np.float64)
    lib().oracle_so3_sums(_p(ni), _p(li), _p(gx), _p(gy), cols, rows, C.byref(pose_curr),
                          C.byref(pose_last), K, _p(sums))
    return sums


def create_matrix3(sums11):
    """CreateMatrix<3,4> (Reduction.cu:6-16) of the float sums."""
    s = np.asarray(sums11, np.float32).astype(np.float64)
    A = np.zeros((3, 3))
    b = np.zeros(3)
    k = 0
    for i in range(3):
        for j in range(i, 4):
            if j == 3:
                b[i] = s[k]
            else:
                A[i, j] = A[j, i] = s[k]
            k += 1
    return A, b


def create_matrix(sums29):
    s = _c(sums29, np.float32)
    A = np.zeros(36, np.float64)
    b = np.zeros(6, np.float64)
    lib().oracle_create_matrix(_p(s), _p(A), _p(b))
    return A.reshape(6, 6), b


def solve6(A, b):
    A = _c(A, np.float64)
    b = _c(b, np.float64)
    x = np.zeros(6, np.float64)
    ok = lib().oracle_cholesky_solve6(_p(A), _p(b), _p(x))
    return x, bool(ok)


def se3_exp(xi):
    xi = _c(xi, np.float64)
    T = np.zeros(16, np.float64)
    lib().oracle_se3_exp(_p(xi), _p(T))
    return T.reshape(4, 4)


def se3_update(A, b, delta, last_pose):
    A = _c(A, np.float64)
    b = _c(b, np.float64)
    delta = _c(delta, np.float64).copy()
    lp = _c(last_pose, np.float64)
    nxt = np.zeros(16, np.float64)
    x = np.zeros(6, np.float64)
    lib().oracle_se3_update(_p(A), _p(b), _p(delta), _p(lp), _p(nxt), _p(x))
    return delta.reshape(4, 4), nxt.reshape(4, 4), x


def euler_sin_norm(Ta, Tb):
    a = _c(Ta, np.float64)
    b = _c(Tb, np.float64)
    rn = C.c_double()
    tn = C.c_double()
    lib().oracle_euler_sin_norm(_p(a), _p(b), C.byref(rn), C.byref(tn))
    return rn.value, tn.value


def se3_control_trace(icp_error, icp_count, iters=(10, 5, 3), min_count=(2000, 1000, 100),
                      rgb_error=None, rgb_count=None, w=1e-4):
    """ComputeSE3's loop decisions on a stream of per-iteration results:
    (final pose marker, failed, updates applied, iterations run, final ICP
    error). Marker 0 is the initial pose, k + 1 the pose after iteration k."""
    e = _c(icp_error, np.float32)
    c = _c(icp_count, np.int32)
    it = _c(iters, np.int32)
    mc = _c(min_count, np.int32)
    re = _c(rgb_error, np.float32) if rgb_error is not None else None
    rc = _c(rgb_count, np.int32) if rgb_count is not None else None
    out = np.zeros(4, np.int32)
    err = C.c_float()
    if lib().oracle_se3_control_trace(_p(it), _p(mc), _p(e), _p(c), _p(re), _p(rc), w, len(e),
                                      _p(out), C.byref(err)) != 0:
        raise ValueError("stream shorter than the loop")
    return int(out[0]), int(out[1]), int(out[2]), int(out[3]), err.value


def se3_gate(icp_error, thresh, pose0, pose):
    return bool(lib().oracle_se3_gate(icp_error, thresh, _p(_c(pose0, np.float64)),
                                      _p(_c(pose, np.float64))))


def measured_pyramid(depth_u16, fx, fy, cx, cy, scale=1000.0, cutoff=3.0):
    """The three (vmap, nmap) levels of a frame as Frame::FillImages makes
    them (Core/Frame.cc:68-92): bilateral filter, PyrDown of the filtered
    depth, vertex and normal maps with K / 2^l."""
    _, d = filter_depth(depth_u16, scale, cutoff)
    out = []
    for lvl in range(3):
        if lvl:
            d = pyrdown_f32(d)
        s = float(np.float32(1 << lvl))
        v = compute_vmap(d, float(np.float32(fx) / np.float32(s)), float(np.float32(fy) / np.float32(s)),
                         float(np.float32(cx) / np.float32(s)), float(np.float32(cy) / np.float32(s)),
                         cutoff)
        out.append((v, compute_nmap(v)))
    return out


def model_pyramid(vmap0, nmap0):
    """The model maps at three levels: the raycast's level 0 and ResizeMap
    below it (Frame::ResizeImages, Core/Frame.cc:94-98)."""
    out = [(np.ascontiguousarray(vmap0, np.float32), np.ascontiguousarray(nmap0, np.float32))]
    for _ in range(2):
        out.append(resize_map(*out[-1]))
    return out


def track_frame(cfg, next_maps, last_maps, last_pose, fp64=True):
    """One ComputeSE3 (point-to-plane) on given pyramids: next_maps and
    last_maps are lists of three (vmap, nmap) float32 pairs (levels 0-2).
    Returns (pose after tracking, tracked)."""
    keep = []

    def arr(maps, k):
        a = (C.c_void_p * 3)()
        for lvl in range(3):
            x = _c(maps[lvl][k], np.float32)
            keep.append(x)
            a[lvl] = x.ctypes.data
        return a
    nv, nn, lv, ln = arr(next_maps, 0), arr(next_maps, 1), arr(last_maps, 0), arr(last_maps, 1)
    lp = _c(last_pose, np.float64)
    out = np.zeros(16, np.float64)
    res = FrameResult()
    ok = lib().oracle_track_frame(C.byref(cfg), 1 if fp64 else 0, C.cast(nv, _P),
                                  C.cast(nn, _P), C.cast(lv, _P), C.cast(ln, _P), _p(lp),
                                  _p(out), C.byref(res))
    return out.reshape(4, 4), bool(ok)


class OracleMap:
    def __init__(self, cfg: MapConfig, fp64=False):
        self.cfg = cfg
        self.fp64 = fp64
        self.h = lib().oracle_map_create(C.byref(cfg), 1 if fp64 else 0)

    def __del__(self):
        try:
            if self.h:
                lib().oracle_map_destroy(self.h)
        except Exception:
            pass

    def reset(self):
        lib().oracle_map_reset(self.h)

    def hash(self, x, y, z):
        return lib().oracle_map_hash(self.h, x, y, z)

    def fuse(self, depth, rgb, nmap, mask, num_detection, pose, K):
        d = _c(depth, np.float32)
        rows, cols = d.shape
        r = _c(rgb, np.uint8)
        n = _c(nmap, np.float32) if nmap is not None else None
        m = _c(mask, np.uint8) if mask is not None else None
        return lib().oracle_fuse_map_color(self.h, _p(d), _p(r), _p(n), _p(m), cols, rows,
                                           num_detection, C.byref(pose), K)

    def fuse_fp64(self, depth, rgb, nmap, mask, num_detection, T, K4):
        """FP64 map: FuseMapColor at the 4x4 double pose T, K4 = (fx, fy, cx, cy)."""
        d = _c(depth, np.float64)
        rows, cols = d.shape
        r = _c(rgb, np.uint8)
        n = _c(nmap, np.float64) if nmap is not None else None
        m = _c(mask, np.uint8) if mask is not None else None
        t = _c(T, np.float64)
        k = _c(K4, np.float64)
        return lib().oracle_fuse_map_color_fp64(self.h, _p(d), _p(r), _p(n), _p(m), cols, rows,
                                                num_detection, _p(t), _p(k))

    def clone_fp64(self):
        """An FP64 map with exactly this FP32 map's state."""
        o = OracleMap.__new__(OracleMap)
        o.cfg = self.cfg
        o.fp64 = True
        o.h = lib().oracle_map_clone_fp64(self.h)
        return o

    def raycast_fp64(self, cols, rows, zmin, zmax, T, invfx, invfy, cx, cy):
        """FP64 map only: the raycast at the 4x4 double pose T."""
        v = np.empty((rows, cols, 4), np.float64)
        n = np.empty((rows, cols, 4), np.float64)
        zmin = _c(zmin, np.float32)
        zmax = _c(zmax, np.float32)
        T = _c(T, np.float64)
        lib().oracle_raycast_fp64(self.h, _p(v), _p(n), cols, rows, _p(zmin), _p(zmax),
                                  zmin.shape[1], _p(T), invfx, invfy, cx, cy)
        return v, n

    def get_sdf(self, slot, idx):
        return lib().oracle_map_get_sdf(self.h, slot, idx)

    def set_sdf(self, slot, idx, v):
        lib().oracle_map_set_sdf(self.h, slot, idx, v)

    def unify_neighbor(self):
        lib().oracle_unify_neighbor(self.h)

    def set_label(self, slot, idx, weight, label, count, label2=0, count2=0):
        lib().oracle_map_set_label(self.h, slot, idx, weight, label, count, label2, count2)

    def mesh(self, tables, max_tris):
        """MeshScene with the given (edge, tri, nvert) tables: (vertices,
        normals, colours) of the triangles written, and the number found."""
        edge, tri, nvert = (np.ascontiguousarray(tables[0], np.int32),
                            np.ascontiguousarray(tables[1], np.int8),
                            np.ascontiguousarray(tables[2], np.int32))
        v = np.zeros((max_tris, 3, 3), np.float32)
        n = np.zeros((max_tris, 3, 3), np.float32)
        c = np.zeros((max_tris, 3, 3), np.uint8)
        tot = np.zeros(1, np.uint32)
        k = lib().oracle_mesh_scene(self.h, _p(edge), _p(tri), _p(nvert), max_tris, _p(v), _p(n),
                                    _p(c), _p(tot))
        return v[:k], n[:k], c[:k], int(tot[0])

    def visibility(self, pose, cols, rows, K, dmax, dmin):
        return lib().oracle_check_visibility(self.h, C.byref(pose), cols, rows, K, dmax, dmin)

    def rendering_blocks(self, zcols, zrows, dmax, dmin, pose, K):
        zmin = np.empty((zrows, zcols), np.float32)
        zmax = np.empty((zrows, zcols), np.float32)
        any_ = lib().oracle_create_rendering_blocks(self.h, _p(zmin), _p(zmax), zcols, zrows,
                                                    dmax, dmin, C.byref(pose), K)
        return zmin, zmax, bool(any_)

    def raycast(self, cols, rows, zmin, zmax, pose, K):
        v = np.empty((rows, cols, 4), np.float32)
        n = np.empty((rows, cols, 4), np.float32)
        zmin = _c(zmin, np.float32)
        zmax = _c(zmax, np.float32)
        lib().oracle_raycast(self.h, _p(v), _p(n), cols, rows, _p(zmin), _p(zmax),
                             zmin.shape[1], C.byref(pose), K)
        return v, n

    def stats(self):
        s = MapStats()
        lib().oracle_map_get_stats(self.h, C.byref(s))
        return {n: getattr(s, n) for n, _ in MapStats._fields_}

    def export(self):
        cfg = self.cfg
        table = np.zeros(2 * cfg.num_buckets, HASH_ENTRY)
        excess = np.zeros(cfg.num_blocks, HASH_ENTRY)
        blocks = np.zeros(cfg.num_blocks, BLOCK)
        geom = np.zeros(cfg.num_blocks * 512, VOXEL_GEOM)
        sem = np.zeros(cfg.num_blocks * 512, VOXEL_SEM)
        lib().oracle_map_export(self.h, _p(table), _p(excess), _p(blocks), _p(geom), _p(sem))
        return dict(table=table, excess=excess, blocks=blocks, geom=geom, sem=sem)

    def visible(self):
        n = lib().oracle_map_visible(self.h, None, 0)
        out = np.zeros(n, BLOCK)
        if n:
            lib().oracle_map_visible(self.h, _p(out), n)
        return out

    def counters(self):
        v = np.zeros(8, np.uint64)
        lib().oracle_map_counters(self.h, _p(v))
        names = ("alloc_probes", "ray_lookups", "march_steps", "ray_norange", "ray_back",
                 "ray_nocross", "ray_nonormal", "ray_hit")
        return {k: int(x) for k, x in zip(names, v)}


class OracleSystem:
    def __init__(self, cfg: SystemConfig, fp64=False):
        self.cfg = cfg
        self.h = lib().oracle_system_create(C.byref(cfg), 1 if fp64 else 0)

    def __del__(self):
        try:
            if self.h:
                lib().oracle_system_destroy(self.h)
        except Exception:
            pass

    def grab(self, depth, rgb, labels=None):
        d = _c(depth, np.uint16)
        r = _c(rgb, np.uint8)
        lab = _c(labels, np.uint8) if labels is not None else None
        res = FrameResult()
        lib().oracle_system_grab(self.h, _p(d), _p(r), _p(lab), 1 if labels is not None else 0,
                                 C.byref(res))
        return dict(frame_index=res.frame_index, tracked=res.tracked, icp_count=res.icp_count,
                    iterations=res.iterations, icp_error=res.icp_error,
                    visible_blocks=res.visible_blocks, allocated_blocks=res.allocated_blocks,
                    pose=np.array(list(res.pose)).reshape(4, 4),
                    rgb_error=res.rgb_error, rgb_count=res.rgb_count)

    def work(self):
        v = np.zeros(8, np.uint64)
        lib().oracle_system_work(self.h, _p(v))
        names = ("raycasts", "rays", "march_steps", "interp_reads", "ray_lookups",
                 "alloc_probes", "updated_voxels", "visible_blocks")
        return {k: int(x) for k, x in zip(names, v)}

    def model(self, level=0):
        cols = self.cfg.cols >> level
        rows = self.cfg.rows >> level
        v = np.empty((rows, cols, 4), np.float32)
        n = np.empty((rows, cols, 4), np.float32)
        lib().oracle_system_model(self.h, level, _p(v), _p(n))
        return v, n
