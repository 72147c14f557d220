"""Pins of the oracle's map functions that do not come from its own code.

- Allocation (FuseMap.cu:92-132, DeviceMap.cu:169-180): every valid pixel's
  accumulated band samples are enumerated in numpy float32, in the
  reference's operation order; the allocated set, and its heap-slot order
  (first request in pixel-major, then step order), must equal the oracle's.
- Visibility (FuseMap.cu:46-90,134-168): a numpy brute force of the
  8-corner test over every allocated block; the visible list must be the
  passing blocks in heap-slot order.
- Rendering prepass (RenderScene.cu:54-190): the reference's loops are
  replayed literally (projectBlock, the tile list of
  createRenderingBlockList with its x < ny bound and unshifted lower-right
  corners, fillBlocks' per-thread guards); the oracle restates them in closed
  form.
- Colour narrowing (VectorMath.h:16-18, FuseMap.cu:647-655): a blend that
  leaves [0, 255] wraps modulo 256 after truncation.

Each test fails on a planted mistake of the kind it names (an off-by-one in
the step count, a non-accumulated band point, a missed corner, a shifted
tile corner, a saturating colour).
"""
import numpy as np
import pytest

from helpers import B, CFG_SMALL, K, map_cfg, synth
from oracle import binding as O

f32 = np.float32


def _pose_arrays(T):
    p = B.pose_f(T)
    R = np.array(list(p.R), f32).reshape(3, 3)
    Rinv = np.array(list(p.Rinv), f32).reshape(3, 3)
    t = np.array(list(p.t), f32)
    return p, R, Rinv, t


def _matvec(M, x, y, z):
    """Matrix3f * float3 (VectorMath.h:368-370): each row's dot product
    a.x*b.x + a.y*b.y + a.z*b.z, left to right, in float32."""
    return [(M[r, 0] * x + M[r, 1] * y) + M[r, 2] * z for r in range(3)]


def _floor_div8(v):
    """voxelPosToBlockPos (DeviceMap.cu:169-180): v - 7 for v < 0, then a
    truncating division by 8."""
    v = np.where(v < 0, v - 7, v)
    return np.fix(v / 8.0).astype(np.int64)


def _float_to_voxel(p):
    """make_int3(float3) (VectorMath.h:36-42): truncate, minus one where the
    truncation is above the value."""
    t = np.trunc(p)
    return np.where(t > p, t - 1, t).astype(np.int64)


def _band_requests(raw, c, T, voxel, trunc):
    """(pixel, step, block) of every CreateBlock call of CreateBlocks, in the
    serial order of the reference's loops: pixel-major, then step."""
    _, R, _, t = _pose_arrays(T)
    rows, cols = raw.shape
    invfx = f32(1.0 / np.float64(f32(c["f"])))
    invfy = invfx
    cx, cy = f32(c["cx"]), f32(c["cy"])
    vinv = f32(1.0 / np.float64(f32(voxel)))
    dmin, dmax = f32(0.1), f32(3.0)
    thresh = f32(f32(trunc) / f32(2))
    ys, xs = np.mgrid[0:rows, 0:cols]
    z = raw.reshape(-1)
    xs, ys = xs.reshape(-1), ys.reshape(-1)
    ok = ~np.isnan(z) & ~(z < dmin) & ~(z > dmax)
    zn = np.maximum(dmin, z - thresh).astype(f32)
    zf = np.minimum(dmax, z + thresh).astype(f32)
    ok &= ~(zn >= zf)
    idx = np.nonzero(ok)[0]
    z, zn, zf, xs, ys = z[idx], zn[idx], zf[idx], xs[idx], ys[idx]

    def unproject(zz):
        px = (zz * (xs.astype(f32) - cx)) * invfx
        py = (zz * (ys.astype(f32) - cy)) * invfy
        w = _matvec(R, px, py, zz)
        return [(w[k] + t[k]) * vinv for k in range(3)]

    near = unproject(zn)
    far = unproject(zf)
    d = [far[k] - near[k] for k in range(3)]
    length = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]).astype(f32)
    nsteps = np.ceil(2.0 * length.astype(np.float64)).astype(np.int64)
    with np.errstate(divide="ignore", invalid="ignore"):
        step = [d[k] / (nsteps - 1).astype(f32) for k in range(3)]
    pts = [near[k].copy() for k in range(3)]
    out = []
    for s in range(int(nsteps.max())):
        live = s < nsteps
        bx = _floor_div8(_float_to_voxel(pts[0]))
        by = _floor_div8(_float_to_voxel(pts[1]))
        bz = _floor_div8(_float_to_voxel(pts[2]))
        for i in np.nonzero(live)[0]:
            out.append((int(idx[i]), s, (int(bx[i]), int(by[i]), int(bz[i]))))
        pts = [(pts[k] + step[k]).astype(f32) for k in range(3)]
    out.sort(key=lambda r: (r[0], r[1]))
    return out


def _fused_small_map(frame=7, frames=None):
    c = CFG_SMALL
    cfg = map_cfg(c, num_blocks=60000)
    m = O.OracleMap(cfg)
    s = synth(c)
    for i in (frames if frames is not None else [frame]):
        d, rgb, lab = s.render(i)
        raw, _ = O.filter_depth(d, 1000.0, 3.0)
        m.fuse(raw, rgb, None, None, 0, B.pose_f(s.pose(i)), K(c))
    return c, cfg, m, s


def test_allocation_is_every_accumulated_band_sample_in_first_request_order():
    c = CFG_SMALL
    cfg = map_cfg(c, num_blocks=60000)
    m = O.OracleMap(cfg)
    s = synth(c)
    T = s.pose(7)  # a rotated and translated camera
    d, rgb, lab = s.render(7)
    raw, _ = O.filter_depth(d, 1000.0, 3.0)
    m.fuse(raw, rgb, None, None, 0, B.pose_f(T), K(c))
    req = _band_requests(raw, c, T, c["voxel"], c["trunc"])
    order = []
    seen = set()
    for _, _, b in req:
        if b not in seen:
            seen.add(b)
            order.append(b)
    n = m.stats()["allocated_blocks"]
    blocks = m.export()["blocks"][:n]
    got = list(zip(blocks["x"].tolist(), blocks["y"].tolist(), blocks["z"].tolist()))
    assert n == len(order), (n, len(order))
    assert got == order
    # the band has several samples per pixel: the step count matters
    steps = np.bincount([r[0] for r in req])
    assert steps[steps > 0].min() >= 2 and steps.max() >= 6


def test_allocation_continues_a_map_with_new_blocks_only():
    """A second frame requests blocks the map holds already: only the new
    ones get heap slots, in first-request order."""
    c = CFG_SMALL
    cfg = map_cfg(c, num_blocks=60000)
    m = O.OracleMap(cfg)
    s = synth(c)
    have = []
    for i in (0, 9):
        d, rgb, lab = s.render(i)
        raw, _ = O.filter_depth(d, 1000.0, 3.0)
        m.fuse(raw, rgb, None, None, 0, B.pose_f(s.pose(i)), K(c))
        for _, _, b in _band_requests(raw, c, s.pose(i), c["voxel"], c["trunc"]):
            if b not in have:
                have.append(b)
    n = m.stats()["allocated_blocks"]
    blocks = m.export()["blocks"][:n]
    assert list(zip(blocks["x"].tolist(), blocks["y"].tolist(), blocks["z"].tolist())) == have


def _visible_brute_force(blocks, T, c, voxel, cols, rows, dmin=0.1, dmax=3.0):
    """CheckBlockVisibility (FuseMap.cu:46-90) of every block: any of the 8
    corners pos*blockWidth + {0, blockWidth}^3 with camera z >= 1e-3, pixel
    0 <= u < cols, 0 <= v < rows and dmin <= z <= dmax."""
    _, _, Rinv, t = _pose_arrays(T)
    bw = f32(f32(voxel) * f32(8))
    fx, fy, cx, cy = f32(c["f"]), f32(c["f"]), f32(c["cx"]), f32(c["cy"])
    pos = [blocks[k].astype(f32) * bw for k in ("x", "y", "z")]
    vis = np.zeros(len(blocks), bool)
    for corner in range(8):
        p = [pos[0] + (bw if corner & 1 else f32(0)), pos[1] + (bw if corner & 2 else f32(0)),
             pos[2] + (bw if corner & 4 else f32(0))]
        q = _matvec(Rinv, p[0] - t[0], p[1] - t[1], p[2] - t[2])
        with np.errstate(divide="ignore", invalid="ignore"):
            u = (fx * q[0]) / q[2] + cx
            v = (fy * q[1]) / q[2] + cy
        ok = ~(q[2] < f32(1e-3)) & (u >= 0) & (v >= 0) & (u < f32(cols)) & (v < f32(rows)) & \
            (q[2] >= f32(dmin)) & (q[2] <= f32(dmax))
        vis |= ok
    return vis


@pytest.mark.parametrize("view", [7, 25])
def test_visibility_is_the_eight_corner_test_in_heap_order(view):
    # blocks of several frames, seen from the pose of `view`: some visible,
    # some not
    c, cfg, m, s = _fused_small_map(frames=[0, 10, 20, 29])
    T = s.pose(view)
    nv = m.visibility(B.pose_f(T), c["cols"], c["rows"], K(c), 3.0, 0.1)
    n = m.stats()["allocated_blocks"]
    blocks = m.export()["blocks"][:n]
    want = _visible_brute_force(blocks, T, c, c["voxel"], c["cols"], c["rows"])
    got = m.visible()
    assert 0 < nv < n
    assert nv == int(want.sum())
    assert np.array_equal(got["slot"], blocks["slot"][want])
    assert np.array_equal(got["x"], blocks["x"][want])


def test_visibility_depth_range_and_image_bounds():
    """One block straight ahead, moved through the depth range and across the
    image edge: visible exactly when a corner passes every bound."""
    c = dict(CFG_SMALL, cols=64, rows=48, f=60.0, cx=31.5, cy=23.5)
    cfg = map_cfg(c, num_buckets=1 << 10, num_blocks=64)
    rng = np.random.default_rng(3)
    for trial in range(40):
        m = O.OracleMap(cfg)
        # one wall pixel in the middle allocates a few blocks around depth z
        z = rng.uniform(0.05, 3.3)
        depth = np.full((c["rows"], c["cols"]), np.nan, f32)
        depth[24, 32] = z
        m.fuse(depth, np.zeros((48, 64, 3), np.uint8), None, None, 0, B.pose_f(np.eye(4)), K(c))
        n = m.stats()["allocated_blocks"]
        if n == 0:
            continue
        T = np.eye(4)
        T[:3, 3] = rng.uniform(-0.4, 0.4, 3)
        nv = m.visibility(B.pose_f(T), 64, 48, K(c), 3.0, 0.1)
        blocks = m.export()["blocks"][:n]
        want = _visible_brute_force(blocks, T, c, c["voxel"], 64, 48)
        assert nv == int(want.sum()), trial
        assert np.array_equal(m.visible()["slot"], blocks["slot"][want]), trial


def _prepass_replay(blocks, T, c, voxel, zc, zr, dmax=3.0, dmin=0.1):
    """CreateRenderingBlocks replayed loop by loop: projectBlock
    (RenderScene.cu:54-112), the tile list of createRenderingBlockList
    (:114-132) over x < ny and y < ny with unshifted lower-right corners, and
    fillBlocks (:167-190) for 16 x 16 threads per tile. zmin starts at 100,
    zmax at 0."""
    _, _, Rinv, t = _pose_arrays(T)
    fx, fy, cx, cy = f32(c["f"]), f32(c["f"]), f32(c["cx"]), f32(c["cy"])
    vs = f32(voxel)
    zmin = np.full((zr, zc), f32(100.0), f32)
    zmax = np.zeros((zr, zc), f32)
    any_tile = False
    for b in blocks:
        ul = [zc, zr]
        lr = [-1, -1]
        zr_ = [f32(dmax), f32(dmin)]
        for corner in range(8):
            tx = int(b["x"]) + (1 if corner & 1 else 0)
            ty = int(b["y"]) + (1 if corner & 2 else 0)
            tz = int(b["z"]) + (1 if corner & 4 else 0)
            p = [f32(tx * 8) * vs, f32(ty * 8) * vs, f32(tz * 8) * vs]
            q = _matvec(Rinv, f32(p[0] - t[0]), f32(p[1] - t[1]), f32(p[2] - t[2]))
            q = [f32(v) for v in q]
            if float(q[2]) < 2e-1:
                continue
            px = f32(f32(f32(fx * q[0]) / q[2]) + cx) / f32(8)
            py = f32(f32(f32(fy * q[1]) / q[2]) + cy) / f32(8)
            if ul[0] > np.floor(px):
                ul[0] = int(np.floor(px))
            if lr[0] < np.ceil(px):
                lr[0] = int(np.ceil(px))
            if ul[1] > np.floor(py):
                ul[1] = int(np.floor(py))
            if lr[1] < np.ceil(py):
                lr[1] = int(np.ceil(py))
            if zr_[0] > q[2]:
                zr_[0] = q[2]
            if zr_[1] < q[2]:
                zr_[1] = q[2]
        ul = [max(ul[0], 0), max(ul[1], 0)]
        lr = [min(lr[0], zc - 1), min(lr[1], zr - 1)]
        if ul[0] > lr[0] or ul[1] > lr[1]:
            continue
        if zr_[0] < f32(dmin):
            zr_[0] = f32(dmin)
        if zr_[1] < f32(dmin):
            continue
        dx = f32(lr[0]) - f32(ul[0]) + f32(1)
        dy = f32(lr[1]) - f32(ul[1]) + f32(1)
        ny = int(np.ceil(dy / f32(16)))
        _ = int(np.ceil(dx / f32(16)))  # nx: computed, never used as a bound
        for y in range(ny):
            for x in range(ny):
                bul = (ul[0] + x * 16, ul[1] + y * 16)
                blr = [ul[0] + 16, ul[1] + 16]
                if blr[0] > lr[0]:
                    blr[0] = lr[0]
                if blr[1] > lr[1]:
                    blr[1] = lr[1]
                for ty in range(16):
                    ypos = bul[1] + ty
                    if ypos > blr[1] or ypos >= zr:
                        continue
                    for tx in range(16):
                        xpos = bul[0] + tx
                        if xpos > blr[0] or xpos >= zc:
                            continue
                        zmin[ypos, xpos] = min(zmin[ypos, xpos], zr_[0])
                        zmax[ypos, xpos] = max(zmax[ypos, xpos], zr_[1])
                        any_tile = True
    return zmin, zmax, any_tile


def test_prepass_replay_of_a_fused_scene():
    c, cfg, m, s = _fused_small_map(frames=[0, 6, 12])
    T = s.pose(12)
    m.visibility(B.pose_f(T), c["cols"], c["rows"], K(c), 3.0, 0.1)
    zc, zr = (c["cols"] + 7) // 8, (c["rows"] + 7) // 8
    zmin, zmax, any_ = m.rendering_blocks(zc, zr, 3.0, 0.1, B.pose_f(T), K(c))
    vis = m.visible()
    wmin, wmax, wany = _prepass_replay(vis, T, c, c["voxel"], zc, zr)
    assert any_ and wany
    assert np.array_equal(zmin.view(np.uint32), wmin.view(np.uint32))
    assert np.array_equal(zmax.view(np.uint32), wmax.view(np.uint32))


def test_prepass_replay_where_the_tile_quirks_act():
    """Blocks of 0.4 m (5 cm voxels) a metre away span more than 16
    subsampled pixels, so the tile list's x < ny bound and its unshifted
    lower-right corners decide which pixels get a range: of a box wider than
    16 pixels only columns ul..ul+16 are filled."""
    c = dict(CFG_SMALL, cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.05,
             trunc=0.15)
    cfg = map_cfg(c, num_buckets=1 << 10, num_blocks=500)
    m = O.OracleMap(cfg)
    depth = np.full((480, 640), 1.2, f32)
    rgb = np.zeros((480, 640, 3), np.uint8)
    m.fuse(depth, rgb, None, None, 0, B.pose_f(np.eye(4)), K(c))
    T = np.eye(4)
    T[:3, 3] = [0.05, -0.03, 0.0]
    nv = m.visibility(B.pose_f(T), 640, 480, K(c), 3.0, 0.1)
    assert nv > 0
    zmin, zmax, any_ = m.rendering_blocks(80, 60, 3.0, 0.1, B.pose_f(T), K(c))
    wmin, wmax, wany = _prepass_replay(m.visible(), T, c, c["voxel"], 80, 60)
    assert any_ and wany
    assert np.array_equal(zmin.view(np.uint32), wmin.view(np.uint32))
    assert np.array_equal(zmax.view(np.uint32), wmax.view(np.uint32))
    # the quirk is active: some block's box is wider than 17 columns, and a
    # column inside its box but past ul + 16 gets nothing from that block
    big = 0
    for b in m.visible():
        xs = []
        for corner in range(8):
            p = np.array([(int(b["x"]) + (corner & 1)) * 0.4 - T[0, 3],
                          (int(b["y"]) + ((corner >> 1) & 1)) * 0.4 - T[1, 3],
                          (int(b["z"]) + ((corner >> 2) & 1)) * 0.4 - T[2, 3]])
            if p[2] >= 0.2:
                xs.append((525.0 * p[0] / p[2] + 319.5) / 8)
        if xs and np.ceil(max(xs)) - np.floor(min(xs)) > 17:
            big += 1
    assert big > 0


def test_colour_blend_narrows_modulo_256():
    """A voxel labelled 2 (palette colour (2, 254, 62), FuseMap.cu:18,570-572)
    seen once more as background with a normal facing the camera at 2.89 m:
    w = n . normalised((p, 1)) is negative, so the blend
    (0.2 w val + (1 - 0.2 w) old) leaves [0, 255] in two channels, and the
    (int) then uchar narrowing (VectorMath.h:16-18) wraps them."""
    c = dict(cols=64, rows=48, f=60.0, cx=31.5, cy=23.5)
    cfg = map_cfg(dict(CFG_SMALL, **c), num_buckets=1 << 12, num_blocks=4000)
    m = O.OracleMap(cfg)
    D = 2.9
    depth = np.full((48, 64), D, f32)
    nmap = np.zeros((48, 64, 4), f32)
    nmap[..., 2] = -1.0
    nmap[..., 3] = 1.0
    k = B.intrinsics(60.0, 60.0, 31.5, 23.5)
    rgb = np.zeros((48, 64, 3), np.uint8)
    rgb[..., 0] = 255
    rgb[..., 2] = 255
    lab2 = np.full((48, 64), 2, np.uint8)
    lab0 = np.zeros((48, 64), np.uint8)
    m.fuse(depth, rgb, nmap, lab2, 1, B.pose_f(np.eye(4)), k)
    m.fuse(depth, rgb, nmap, lab0, 1, B.pose_f(np.eye(4)), k)
    ex = m.export()
    vz = 289  # corner z = 2.89 m, 1 cm in front of the wall
    blocks = ex["blocks"]
    hit = np.nonzero((blocks["x"] == 0) & (blocks["y"] == 0) & (blocks["z"] == vz // 8) &
                     (blocks["slot"] >= 0))[0]
    assert len(hit) == 1
    idx = hit[0] * 512 + (vz % 8) * 64
    g, s = ex["geom"][idx], ex["sem"][idx]
    assert (s["label"], s["count"], s["label2"], s["count2"]) == (0, 1, 2, 1)
    assert (s["r2"], s["g2"], s["b2"]) == (2, 254, 62)
    # the blend in exact arithmetic
    p = np.array([0.0, 0.0, vz * 0.01, 1.0])
    w = np.dot([0.0, 0.0, -1.0, 1.0], p / np.linalg.norm(p))
    assert w < -0.6
    val = np.array([255.0, 0.0, 255.0]) / 255.0
    old = np.array([2.0, 254.0, 62.0]) / 255.0
    res = (w * 0.2 * val + (1 - w * 0.2) * old) * 255.0
    assert np.all(np.abs(res - np.round(res)) > 0.05)  # no truncation edge case
    want = np.trunc(res).astype(np.int64) % 256
    assert res[0] < 0 and res[1] > 255  # both sides of the range
    assert (g["r"], g["g"], g["b"]) == tuple(int(x) for x in want)
    assert tuple(int(x) for x in want) == (227, 29, 38)
