"""Pins of the oracle (CPU restatement) against facts that do not come from
its own code: closed forms, textbook routines (scipy's matrix exponential,
numpy's linear solve, a plain 5x5 convolution), hand-traced sequences of the
label state machine, and invariants of the hash table. The reference holds
no golden vectors or tests for this path, so these are the oracle's anchors.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg
import scipy.ndimage

from helpers import B, CFG_SMALL, K, map_cfg, synth
from oracle import binding as O


# ---------------------------------------------------------------- exp ------
def test_expf_within_two_ulp():
    xs = np.concatenate([np.linspace(-103.0, 88.0, 20001, dtype=np.float32),
                         np.float32([-0.0, 0.0, -1e-8, 1e-8, -0.5, -18.0, -87.3])])
    got = np.array([O.expf(float(x)) for x in xs], np.float32)
    ref = np.exp(xs.astype(np.float64))
    normal = ref > np.finfo(np.float32).tiny
    ulp = np.abs(got[normal].astype(np.float64) - ref[normal]) / np.spacing(
        ref[normal].astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 2.0, ulp.max()
    sub = ~normal
    assert np.all(np.abs(got[sub] - ref[sub]) <= 2 * 1.4e-45)


# --------------------------------------------------------------- hash ------
def test_hash_known_values():
    cfg = map_cfg(CFG_SMALL, num_buckets=1000000)
    m = O.OracleMap(cfg)
    # (1,0,0): 73856093 mod 1e6
    assert m.hash(1, 0, 0) == 856093
    # (-1,0,0): the int XOR converts to unsigned before `% NumBuckets`
    # (C++ usual arithmetic conversions), giving (2^32 - 73856093) mod 1e6 =
    # 111203, not the signed remainder + N = 143907.
    assert m.hash(-1, 0, 0) == 111203
    assert m.hash(0, -1, 0) == (2**32 - 19349669) % 1000000
    # a product that wraps: 100 * 83492791 overflows int32
    assert m.hash(0, 0, 100) == ((100 * 83492791) % 2**32) % 1000000


def test_hash_power_of_two_is_low_bits():
    cfg = map_cfg(CFG_SMALL, num_buckets=1 << 18)
    m = O.OracleMap(cfg)
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(-3000, 3000, size=(200, 3)):
        h = ((int(x) * 73856093) ^ (int(y) * 19349669) ^ (int(z) * 83492791)) & 0xFFFFFFFF
        assert m.hash(int(x), int(y), int(z)) == h & ((1 << 18) - 1)


# ------------------------------------------------------------- pyramid -----
def test_pyrdown_constant_fixed_point_and_interior_convolution():
    c = np.full((40, 52), 1.75, np.float32)
    assert np.all(O.pyrdown_f32(c) == np.float32(1.75))
    assert np.all(O.pyrdown_u8(np.full((40, 52), 201, np.uint8)) == 201)
    rng = np.random.default_rng(1)
    img = rng.uniform(0.5, 3.0, size=(40, 52)).astype(np.float32)
    out = O.pyrdown_f32(img)
    k1 = np.array([1, 4, 6, 4, 1], np.float64)
    conv = scipy.ndimage.convolve(img.astype(np.float64), np.outer(k1, k1) / 256.0)
    ref = conv[::2, ::2][: out.shape[0], : out.shape[1]]
    # interior: the full window lies inside [0, rows-1) x [0, cols-1)
    assert np.allclose(out[1:-2, 1:-2], ref[1:-2, 1:-2], rtol=2e-6, atol=0)


def test_pyrdown_skips_nan_taps():
    img = np.full((20, 20), 2.0, np.float32)
    img[8:12, 8:12] = np.nan
    out = O.pyrdown_f32(img)
    assert np.all(out == 2.0)  # NaN taps dropped from both sums
    img[:] = np.nan
    assert np.all(np.isnan(O.pyrdown_f32(img)))  # 0/0 when every tap is NaN


def test_filter_depth_constant_and_last_column_excluded():
    d = np.full((30, 40), 1500, np.uint16)
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    # depth * (float)(1 / scale), the inverse rounded to float (Pyrdown.cu:81,117)
    z = np.float32(1500) * np.float32(1.0 / 1000.0)
    assert np.all(raw == z)
    assert np.allclose(filt, z, rtol=4e-7, atol=0)  # weighted mean of equal values: a few ulp
    # The window stops before the last column (Pyrdown.cu:90,96): a spike in
    # the last column is invisible to the filtered value of that column.
    d[:, -1] = 2500
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    assert np.all(raw[:, -1] == np.float32(2500) * np.float32(1.0 / 1000.0))
    assert np.allclose(filt[:, -1], z, rtol=4e-7, atol=0)
    # cutoff: raw and filtered are NaN at and beyond it
    d = np.full((10, 10), 3100, np.uint16)
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    assert np.all(np.isnan(raw)) and np.all(np.isnan(filt))


def test_filter_depth_matches_textbook_bilateral_in_interior():
    rng = np.random.default_rng(2)
    d = (1000 + rng.integers(-300, 300, size=(24, 24))).astype(np.uint16)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    z = d.astype(np.float64) / 1000.0
    ref = np.zeros_like(z)
    for y in range(2, 21):
        for x in range(2, 21):
            w_sum = v_sum = 0.0
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    v = z[y + dy, x + dx]
                    w = np.exp(-((dx * dx + dy * dy) / 32.0 + (z[y, x] - v) ** 2 * 2.0))
                    w_sum += w
                    v_sum += w * v
            ref[y, x] = v_sum / w_sum
    assert np.allclose(filt[2:21, 2:21], ref[2:21, 2:21], rtol=2e-6, atol=0)


def test_vmap_nmap_plane_closed_form():
    rows, cols, f, cx, cy = 30, 40, 100.0, 19.5, 14.5
    depth = np.full((rows, cols), 1.2, np.float32)
    depth[0, 0] = 0.05   # too close
    depth[0, 1] = np.nan
    v = O.compute_vmap(depth, f, f, cx, cy, 3.0)
    ys, xs = np.mgrid[0:rows, 0:cols]
    assert np.allclose(v[..., 0][2:, 2:], ((xs - cx) / f * 1.2)[2:, 2:], rtol=1e-6, atol=1e-7)
    assert np.allclose(v[..., 1][2:, 2:], ((ys - cy) / f * 1.2)[2:, 2:], rtol=1e-6, atol=1e-7)
    assert np.isnan(v[0, 0, 0]) and np.isnan(v[0, 1, 0])
    n = O.compute_nmap(v)
    # fronto-parallel plane: cross(right - c, down - c) = +z
    inner = n[2:-1, 2:-1]
    assert np.allclose(inner[..., :3], [0, 0, 1], atol=1e-6)
    assert np.all(np.isnan(n[-1, :, 0])) and np.all(np.isnan(n[:, -1, 0]))
    # tilted plane z = 1 + 0.2 x_cam: normal along (-0.2, 0, 1) up to sign
    zt = (1.0 / (1 - 0.2 * (xs - cx) / f)).astype(np.float32)
    nt = O.compute_nmap(O.compute_vmap(zt, f, f, cx, cy, 3.0))
    ref = np.array([-0.2, 0, 1]) / np.linalg.norm([-0.2, 0, 1])
    dots = np.abs((nt[1:-1, 1:-1, :3] * ref).sum(-1))
    assert dots.min() > 1 - 1e-5


def test_resize_map_normal_length_is_one_over_sqrt2():
    v = np.zeros((8, 8, 4), np.float32)
    v[..., 2] = 1.0
    v[..., 3] = 1.0
    n = np.zeros((8, 8, 4), np.float32)
    n[..., 0] = 0.6
    n[..., 2] = 0.8
    n[..., 3] = 1.0
    vd, nd = O.resize_map(v, n)
    assert np.allclose(np.linalg.norm(nd[..., :3], axis=-1), 1 / np.sqrt(2), atol=1e-7)
    assert np.allclose(nd[..., 3], 1 / np.sqrt(2), atol=1e-7)
    v[3, 3, 0] = np.nan
    vd, _ = O.resize_map(v, n)
    assert np.isnan(vd[1, 1, 0]) and not np.isnan(vd[0, 0, 0])


def test_intensity_known_values():
    rgb = np.zeros((1, 4, 3), np.uint8)
    rgb[0, 0] = (255, 0, 0)
    rgb[0, 1] = (0, 255, 0)
    rgb[0, 2] = (0, 0, 255)
    rgb[0, 3] = (10, 20, 30)
    g = O.image_to_intensity(rgb)[0]
    # weights 0.7154 on the first channel, 0.2125 on the second, 0.0721 on the third
    assert list(g) == [182, 54, 18, int(0.2125 * 20 + 0.7154 * 10 + 0.0721 * 30)]


def test_derivative_ramps():
    xs = np.arange(20)
    img = np.tile((3 * xs).astype(np.uint8), (15, 1))
    dx, dy = O.derivative_image(img)
    assert np.all(dx[1:-1, 1:-1] == 24) and np.all(dy[1:-1, 1:-1] == 0)
    img = np.tile((5 * np.arange(15)).astype(np.uint8)[:, None], (1, 20))
    dx, dy = O.derivative_image(img)
    assert np.all(dy[1:-1, 1:-1] == 40) and np.all(dx[1:-1, 1:-1] == 0)


# ----------------------------------------------------------- ICP / solve ---
def _frame_maps(c, i, noise=False):
    d, _, _ = synth(c, noise=noise, dropout=0.0).render(i)
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    v = O.compute_vmap(raw, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    return v, O.compute_nmap(v)


def test_icp_identical_frames_closed_form():
    c = CFG_SMALL
    v, n = _frame_maps(c, 0)
    p = B.pose_f(np.eye(4))
    s = O.icp_sums(v, v, n, n, p, p, K(c))
    A, b = O.create_matrix(s.astype(np.float32))
    valid = (~np.isnan(v[..., 0])) & (v[..., 2] >= 1e-3) & (~np.isnan(n[..., 0]))
    assert s[28] == valid.sum()
    assert s[27] == 0.0 and np.all(b == 0.0)
    assert np.allclose(A, A.T) and np.linalg.eigvalsh(A).min() > -1e-9


def test_icp_plane_is_rank_three():
    rows, cols, f = 60, 80, 80.0
    depth = np.full((rows, cols), 1.0, np.float32)
    v = O.compute_vmap(depth, f, f, 39.5, 29.5, 3.0)
    n = O.compute_nmap(v)
    p = B.pose_f(np.eye(4))
    s = O.icp_sums(v, v, n, n, p, p, B.intrinsics(f, f, 39.5, 29.5))
    A, _ = O.create_matrix(s.astype(np.float32))
    w = np.linalg.eigvalsh(A)
    assert (w > 1e-6 * w.max()).sum() == 3


def test_icp_recovers_known_motion():
    """Gauss-Newton on two noise-free frames of a known relative motion
    converges to it: pins the Jacobian signs, the twist order and the update
    composition next = last * (exp(x) * delta)^-1."""
    c = CFG_SMALL
    v0, n0 = _frame_maps(c, 0)
    v1, n1 = _frame_maps(c, 2)
    gt = synth(c).pose(2)
    last, nxt, delta = np.eye(4), np.eye(4), np.eye(4)
    for _ in range(15):
        s = O.icp_sums(v1, v0, n1, n0, B.pose_f(nxt), B.pose_f(last), K(c))
        A, b = O.create_matrix(s.astype(np.float32))
        delta, nxt, _ = O.se3_update(A, b, delta, last)
    assert np.abs(nxt[:3, 3] - gt[:3, 3]).max() < 3e-4
    assert np.abs(nxt[:3, :3] - gt[:3, :3]).max() < 3e-4


# ------------------------------------------------------------ photometric --
PLANE = dict(cols=96, rows=72, f=90.0, cx=47.5, cy=35.5)


def _plane(rng, zeros=0.0):
    """A textured fronto-parallel plane at 1 m: vertex map, intensity with
    optional zero pixels, and its derivatives."""
    rows, cols = PLANE["rows"], PLANE["cols"]
    v = O.compute_vmap(np.ones((rows, cols), np.float32), PLANE["f"], PLANE["f"], PLANE["cx"],
                       PLANE["cy"], 3.0)
    img = rng.integers(20, 230, size=(rows, cols)).astype(np.uint8)
    if zeros:
        img[rng.random((rows, cols)) < zeros] = 0
    dx, dy = O.derivative_image(img)
    return v, img, dx, dy


def _plane_K():
    return B.intrinsics(PLANE["f"], PLANE["f"], PLANE["cx"], PLANE["cy"])


def _candidates(img, dx, dy, v, shift=0):
    """FindCorresp's tests written out per pixel: border 5, the window rows
    y-2..y+1 and columns x-2..x+1 positive, a non-zero gradient, a valid
    vertex; with a shift of the last frame, the matched pixel x + shift in
    the image and non-zero there."""
    rows, cols = img.shape
    ok = np.zeros((rows, cols), bool)
    for y in range(5, rows - 5):
        for x in range(5, cols - 5):
            u = x + shift
            ok[y, x] = (np.all(img[y - 2:y + 2, x - 2:x + 2] > 0) and
                        (int(dx[y, x]) ** 2 + int(dy[y, x]) ** 2) > 0 and
                        not np.isnan(v[y, x, 0]) and 0 <= u < cols)
    return ok


def _photometric_A(img, dx, dy, v, ok):
    """sum J^T J of the textbook photometric Jacobian
    J = grad(I) . d(pi)/dp . [I | -[p]x] (twist order translation, rotation)."""
    f = PLANE["f"]
    A = np.zeros((6, 6))
    Jsum = np.zeros(6)
    for y, x in zip(*np.nonzero(ok)):
        p = v[y, x, :3].astype(np.float64)
        g = np.array([dx[y, x], dy[y, x]], np.float64) / 9.0
        Jpi = np.array([[f / p[2], 0, -f * p[0] / p[2] ** 2], [0, f / p[2], -f * p[1] / p[2] ** 2]])
        P = np.array([[0, -p[2], p[1]], [p[2], 0, -p[0]], [-p[1], p[0], 0]])
        J = g @ Jpi @ np.hstack([np.eye(3), -P])
        A += np.outer(J, J)
        Jsum += J
    return A, Jsum


# ------------------------------------------- compensated FP32 sums --------
def _k32_rows(n, col0):
    r = np.zeros((n, 29), np.float32)
    r[:, 28] = 1.0
    for i, v in col0.items():
        r[i, 0] = v
    return r


def test_k32_sums_known_answers():
    """2^24 + 1 + 1 is 16777218 exactly; plain FP32 loses both ones (each
    addition ties back to 2^24). Pixels 0, 1, 2 meet in the workgroup's
    butterfly (TwoSum); pixels 0, 256, 512 are one thread's (Kahan)."""
    big = float(2 ** 24)
    for n, idx in ((3, (0, 1, 2)), (513, (0, 256, 512))):
        r = _k32_rows(n, {idx[0]: big, idx[1]: 1.0, idx[2]: 1.0})
        s = O.sums_k32(r)
        assert s[0] == 16777218.0 and s[28] == n
        assert np.all(s[1:28] == 0)
        assert O.sums_k32(r, compensate=False)[0] == big
    # rows whose slot 28 is 0 do not contribute
    r = _k32_rows(40, {3: 5.0, 7: 2.0})
    r[3, 28] = 0.0
    s = O.sums_k32(r)
    assert s[0] == 2.0 and s[28] == 39


def _exact_bound_ratio(rows, s):
    """max over the 29 sums of |s - exact| / (2^-24 |exact| + 2^-40 sum |x|)."""
    import math
    r64 = rows.astype(np.float64)
    worst = 0.0
    for j in range(29):
        col = r64[:, j][rows[:, 28] != 0]
        exact = math.fsum(col)
        bound = 2.0 ** -24 * abs(exact) + 2.0 ** -40 * math.fsum(np.abs(col))
        err = abs(float(s[j]) - exact)
        worst = max(worst, err / bound if bound > 0 else (0.0 if err == 0 else np.inf))
    return worst


@pytest.mark.parametrize("case", ["icp", "cancel"])
def test_k32_sums_within_half_ulp_of_exact(case):
    """The compensated sum is the exact sum of the FP32 inputs rounded once
    (within 2^-40 of the sum of magnitudes), over several workgroups; the same
    order without compensation (a planted mistake) misses that bound."""
    if case == "icp":
        c = CFG_SMALL
        v0, n0 = _frame_maps(c, 0)
        v1, n1 = _frame_maps(c, 3)
        _, rows = O.icp_sums(v1, v0, n1, n0, B.pose_f(synth(c).pose(3)), B.pose_f(np.eye(4)),
                             K(c), rows_out=True)
    else:
        rng = np.random.default_rng(11)
        n = 5 * 1024 + 37
        rows = (rng.standard_normal((n, 29)) * 10.0 ** rng.integers(-3, 4, (n, 29))).astype(
            np.float32)
        rows[:, 28] = (rng.random(n) < 0.9).astype(np.float32)
    assert _exact_bound_ratio(rows, O.sums_k32(rows)) <= 1.0
    assert _exact_bound_ratio(rows, O.sums_k32(rows, compensate=False)) > 1.0


def test_rgb_identical_frames_textbook_jacobian():
    """Zero motion, identical frames with scattered zero intensities: the
    count is the number of pixels passing the candidate tests, every
    difference is 0 (sigma 0, weight 1, b 0), and A is the textbook sum."""
    rng = np.random.default_rng(11)
    v, img, dx, dy = _plane(rng, zeros=0.01)
    p = B.pose_f(np.eye(4))
    s, (cnt, sq), sigma = O.rgb_sums(img, img, v, v, dx, dy, p, p, _plane_K())
    ok = _candidates(img, dx, dy, v)
    assert cnt == ok.sum() == s[28] and sq == 0 and sigma == 0.0
    A, b = O.create_matrix(s.astype(np.float32))
    assert s[27] == 0.0 and np.all(b == 0.0)
    At, _ = _photometric_A(img, dx, dy, v, ok)
    assert np.abs(A - At).max() < 1e-5 * np.abs(At).max()


def test_rgb_constant_offset_weight_and_residual():
    """The last image brighter by 10 everywhere: every difference is -10, so
    sum d^2 = 100 n, sigma = 10, the weight 1 / (sigma + |d|) = 1/20, the
    residual -w d = 0.5: A = A0 / 400, b = sum(J) / 40, sum r^2 = n / 4."""
    rng = np.random.default_rng(12)
    v, img, dx, dy = _plane(rng)
    last = (img.astype(np.int32) + 10).astype(np.uint8)
    p = B.pose_f(np.eye(4))
    s, (cnt, sq), sigma = O.rgb_sums(img, last, v, v, dx, dy, p, p, _plane_K())
    ok = _candidates(img, dx, dy, v)
    n = ok.sum()
    assert cnt == n == s[28] and sq == 100 * n and sigma == 10.0
    A, b = O.create_matrix(s.astype(np.float32))
    At, Jsum = _photometric_A(img, dx, dy, v, ok)
    assert np.abs(A - At / 400).max() < 1e-5 * np.abs(At / 400).max()
    assert np.abs(b - Jsum / 40).max() < 1e-5 * np.abs(Jsum / 40).max()
    assert abs(s[27] - n / 4) < 1e-6 * n


@pytest.mark.parametrize("shift", [3, -2])
def test_rgb_projection_follows_translation(shift):
    """The current camera moved by t along x over a plane at 1 m: pixel x
    sees the last frame's pixel x + f t. With the last image shifted to
    match, every correspondence has difference 0, and the count is the
    candidates whose matched pixel is inside the image."""
    rng = np.random.default_rng(13)
    v, img, dx, dy = _plane(rng)
    cols = PLANE["cols"]
    last = np.zeros_like(img)
    if shift > 0:
        last[:, shift:] = img[:, :cols - shift]
    else:
        last[:, :cols + shift] = img[:, -shift:]
    T = np.eye(4)
    T[0, 3] = shift / PLANE["f"]
    s, (cnt, sq), _ = O.rgb_sums(img, last, v, v, dx, dy, B.pose_f(T), B.pose_f(np.eye(4)),
                                 _plane_K())
    ok = _candidates(img, dx, dy, v, shift)
    assert cnt == ok.sum() and sq == 0 and s[27] == 0.0


def test_rgb_gauss_newton_moves_towards_known_motion():
    """Photometric-only Gauss-Newton on two rendered frames of a known motion
    reduces the pose error and the intensity residual: pins the Jacobian
    signs against the update composition (its nearest-pixel lookup stops
    it at a few millimetres)."""
    c = CFG_SMALL

    def frame(i):
        d, rgb, _ = synth(c, noise=False, dropout=0.0).render(i)
        raw, _ = O.filter_depth(d, 1000.0, 3.0)
        v = O.compute_vmap(raw, c["f"], c["f"], c["cx"], c["cy"], 3.0)
        g = O.image_to_intensity(rgb)
        return (v, g) + tuple(O.derivative_image(g))

    v0, g0, _, _ = frame(0)
    v1, g1, dx1, dy1 = frame(2)
    gt = synth(c).pose(2)
    last, nxt, delta = np.eye(4), np.eye(4), np.eye(4)
    err = lambda T: (np.abs(T[:3, 3] - gt[:3, 3]).max(), np.abs(T[:3, :3] - gt[:3, :3]).max())
    t0, r0 = err(nxt)
    sq = []
    for _ in range(5):
        s, (cnt, q), _ = O.rgb_sums(g1, g0, v1, v0, dx1, dy1, B.pose_f(nxt), B.pose_f(last), K(c))
        sq.append(q / cnt)
        A, b = O.create_matrix(s.astype(np.float32))
        delta, nxt, _ = O.se3_update(A, b, delta, last)
    t1, r1 = err(nxt)
    assert t1 < 0.3 * t0 and r1 < 0.7 * r0
    assert sq[-1] < 0.2 * sq[0]


def _skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _rot_pose(R):
    T = np.eye(4)
    T[:3, :3] = R
    return B.pose_f(T)


def _so3_A_textbook(img_dx, img_dy, rows, cols, f, cx, cy, offset_r=None):
    """sum J^T J (and sum J r) of the rotation-only photometric Jacobian
    J = grad(I) . d(pi)/dp . (-[p]x) at every interior pixel (identity
    rotations: the match is the pixel itself, p its bearing)."""
    A = np.zeros((3, 3))
    Jsum = np.zeros(3)
    for y in range(5, rows - 5):
        for x in range(5, cols - 5):
            p = np.array([(x - cx) / f, (y - cy) / f, 1.0])
            g = np.array([img_dx[y, x], img_dy[y, x]], np.float64) / 9.0
            Jpi = np.array([[f / p[2], 0, -f * p[0] / p[2] ** 2],
                            [0, f / p[2], -f * p[1] / p[2] ** 2]])
            J = g @ Jpi @ (-_skew(p))
            A += np.outer(J, J)
            Jsum += J
    return A, Jsum


def test_so3_identical_frames_textbook_jacobian():
    """Identity rotations on identical images: every interior pixel (5-pixel
    border) matches itself by the rounded projection, r = 0, and A is the
    textbook rotation-only Jacobian summed in FP64."""
    rng = np.random.default_rng(21)
    _, img, dx, dy = _plane(rng)
    rows, cols = img.shape
    s = O.so3_sums(img, img, dx, dy, _rot_pose(np.eye(3)), _rot_pose(np.eye(3)), _plane_K())
    assert s[10] == (rows - 10) * (cols - 10)
    A, b = O.create_matrix3(s)
    assert s[9] == 0.0 and np.all(b == 0.0)
    At, _ = _so3_A_textbook(dx, dy, rows, cols, PLANE["f"], PLANE["cx"], PLANE["cy"])
    assert np.abs(A - At).max() < 1e-5 * np.abs(At).max()


def test_so3_constant_offset_residual():
    """The last image brighter by 10: r = -(I_next - I_last) = 10 at every
    match, so sum r^2 = 100 n and b = 10 sum(J)."""
    rng = np.random.default_rng(22)
    _, img, dx, dy = _plane(rng)
    rows, cols = img.shape
    last = (img.astype(np.int32) + 10).astype(np.uint8)
    s = O.so3_sums(img, last, dx, dy, _rot_pose(np.eye(3)), _rot_pose(np.eye(3)), _plane_K())
    n = (rows - 10) * (cols - 10)
    assert s[10] == n and abs(s[9] - 100 * n) < 1e-9 * n
    _, b = O.create_matrix3(s)
    _, Jsum = _so3_A_textbook(dx, dy, rows, cols, PLANE["f"], PLANE["cx"], PLANE["cy"])
    assert np.abs(b - 10 * Jsum).max() < 1e-5 * np.abs(10 * Jsum).max()


def test_so3_rounded_projection_and_pixel_roles():
    """A yaw of 0.6 pixel (at the centre; up to 0.8 at the edges) moves every
    match one column right under round-to-nearest (truncation would keep it):
    with the last image shifted one column to match, every residual
    -(I_next(u, v) - I_last(x, y)) is 0, and the count loses the one column
    whose match leaves the border."""
    rng = np.random.default_rng(23)
    _, img, dx, dy = _plane(rng)
    rows, cols = img.shape
    Rc = scipy.linalg.expm(_skew([0, -0.6 / PLANE["f"], 0]))
    last = np.zeros_like(img)
    last[:, :-1] = img[:, 1:]
    s = O.so3_sums(img, last, dx, dy, _rot_pose(Rc), _rot_pose(np.eye(3)), _plane_K())
    assert s[10] == (rows - 10) * (cols - 11) and s[9] == 0.0


def test_so3_step_recovers_known_rotation():
    """Images that depend only on the viewing direction (a pure rotation is
    an exact warp): one Gauss-Newton step of ComputeSO3's update
    (delta = exp(x) delta, next = last delta^-1) removes most of a known
    rotation, which pins the Jacobian's sign against the update."""
    rows, cols, f, cx, cy = 120, 160, 131.25, 79.5, 59.5

    def render(R):
        ys, xs = np.mgrid[0:rows, 0:cols].astype(np.float64)
        d = np.stack([(xs - cx) / f, (ys - cy) / f, np.ones_like(xs)], -1) @ R.T
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        i = 128 + 50 * np.sin(7 * d[..., 0]) * np.cos(5 * d[..., 1]) + \
            30 * np.sin(4 * d[..., 1] + 2 * d[..., 0])
        return np.clip(np.round(i), 1, 255).astype(np.uint8)

    Rc = scipy.linalg.expm(_skew([0.01, -0.015, 0.008]))
    last, nxt = render(np.eye(3)), render(Rc)
    dx, dy = O.derivative_image(nxt)
    s = O.so3_sums(nxt, last, dx, dy, _rot_pose(np.eye(3)), _rot_pose(np.eye(3)),
                   B.intrinsics(f, f, cx, cy))
    A, b = O.create_matrix3(s)
    delta = scipy.linalg.expm(_skew(np.linalg.solve(A, b)))
    nxt_rot = np.linalg.inv(delta)
    assert np.abs(nxt_rot - Rc).max() < 0.25 * np.abs(np.eye(3) - Rc).max()


def test_solve_matches_numpy():
    rng = np.random.default_rng(4)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 1e-3 * np.eye(6)
        b = rng.normal(size=6)
        x, ok = O.solve6(A, b)
        assert ok
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("scale", [1e-13, 1e-6, 1e-2, 0.5, 2.0])
def test_se3_exp_matches_matrix_exponential(scale):
    rng = np.random.default_rng(5)
    for _ in range(10):
        xi = rng.normal(size=6) * scale
        v, w = xi[:3], xi[3:]
        W = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
        tw = np.zeros((4, 4))
        tw[:3, :3] = W
        tw[:3, 3] = v
        assert np.allclose(O.se3_exp(xi), scipy.linalg.expm(tw), rtol=0, atol=1e-12)


def test_euler_sine_norm_small_rotations():
    for ax in range(3):
        w = np.zeros(3)
        w[ax] = 0.05
        T = O.se3_exp(np.concatenate([[0.01, -0.02, 0.03], w]))
        rn, tn = O.euler_sin_norm(np.eye(4), T)
        assert abs(rn - np.sin(0.05)) < 1e-12
        assert abs(tn - np.linalg.norm(T[:3, 3])) < 1e-15


# ---------------------------------------------------------------- map ------
WALL = dict(cols=64, rows=48, f=60.0, cx=31.5, cy=23.5, voxel=0.01, trunc=0.03,
            buckets=1 << 12, blocks=4000)


def _wall_fuse(m, depth_m, label, frames=1, rgb_value=100):
    rows, cols = WALL["rows"], WALL["cols"]
    depth = np.full((rows, cols), depth_m, np.float32)
    rgb = np.full((rows, cols, 3), rgb_value, np.uint8)
    v = O.compute_vmap(depth, WALL["f"], WALL["f"], WALL["cx"], WALL["cy"], 3.0)
    n = O.compute_nmap(v)
    n[-1, :] = n[-2, :]
    n[:, -1] = n[:, -2]
    lab = np.full((rows, cols), label, np.uint8)
    k = B.intrinsics(WALL["f"], WALL["f"], WALL["cx"], WALL["cy"])
    for _ in range(frames):
        m.fuse(depth, rgb, n, lab, 1, B.pose_f(np.eye(4)), k)


def _voxel(m_export, vx, vy, vz):
    bx, by, bz = vx // 8, vy // 8, vz // 8
    blocks = m_export["blocks"]
    hit = np.nonzero((blocks["x"] == bx) & (blocks["y"] == by) & (blocks["z"] == bz) &
                     (blocks["slot"] >= 0))[0]
    assert len(hit) == 1
    idx = hit[0] * 512 + (vz % 8) * 64 + (vy % 8) * 8 + (vx % 8)
    return m_export["geom"][idx], m_export["sem"][idx]


def test_tsdf_fronto_parallel_wall_closed_form():
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    D = 1.0
    _wall_fuse(m, D, 0, frames=3)
    ex = m.export()
    # voxel corners on the optical axis, (0, 0, z), in the allocated block
    # z = 96..103 around the wall (the +-T/2 band is 98.5..101.5 voxels)
    for vz in range(96, 104):
        g, _ = _voxel(ex, 0, 0, vz)
        z = np.float32(vz) * np.float32(0.01)
        sdf = np.float32(D) - z
        if sdf >= -0.03:
            assert g["weight"] == 3
            assert abs(g["sdf"] - min(1.0, sdf / 0.03)) < 2e-6, (vz, g["sdf"])
            assert (g["r"], g["g"], g["b"]) == (100, 100, 100)
        else:
            assert g["weight"] == 0 and np.isnan(g["sdf"])


def test_label_state_machine_hand_traced():
    """One voxel in front of a wall observed with label streams whose
    outcomes were traced by hand through FuseMap.cu:575-662."""
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    expected = [  # label of the frame -> (label, count, label2, count2) after it
        (3, (3, 1, 0, 0)),   # new voxel, backup takes 3, swap (count 0 < 2*1, label 0)
        (3, (3, 2, 0, 0)),
        (3, (3, 3, 0, 0)),
        (5, (3, 3, 5, 1)),   # backup 5, no swap (3 < 1 false)
        (5, (3, 3, 5, 2)),
        (5, (3, 3, 5, 3)),
        (5, (5, 4, 3, 3)),   # 3 < 4 with label 3 != 0 -> swap
        (0, (0, 1, 5, 4)),   # background, count 4 < 7 -> demote to backup
        (0, (0, 2, 5, 4)),
        (5, (5, 5, 0, 2)),   # backup hit; 2 < 2*5 with label 0 -> swap
        (7, (5, 5, 0, 2)),   # count2 2 < 7 -> backup 7 ...
    ]
    # the last row: backup replaced by 7 (count2 = 1), no swap (5 < 1 false)
    expected[-1] = (7, (5, 5, 7, 1))
    for label, want in expected:
        _wall_fuse(m, 1.0, label)
        _, s = _voxel(m.export(), 0, 0, 99)
        assert (s["label"], s["count"], s["label2"], s["count2"]) == want, (label, s)


def test_label_count_saturates_at_cap():
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"],
                  max_label_count=255)
    m = O.OracleMap(cfg)
    for _ in range(260):
        _wall_fuse(m, 1.0, 4)
    _, s = _voxel(m.export(), 0, 0, 99)
    assert s["label"] == 4 and s["count"] == 255


def test_colour_running_mean_constant():
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    m.fuse(np.full((48, 64), 1.0, np.float32), np.full((48, 64, 3), 77, np.uint8), None, None, 0,
           B.pose_f(np.eye(4)), B.intrinsics(60.0, 60.0, 31.5, 23.5))
    for _ in range(9):
        m.fuse(np.full((48, 64), 1.0, np.float32), np.full((48, 64, 3), 77, np.uint8), None,
               None, 0, B.pose_f(np.eye(4)), B.intrinsics(60.0, 60.0, 31.5, 23.5))
    g, _ = _voxel(m.export(), 0, 0, 99)
    assert g["weight"] == 10 and (g["r"], g["g"], g["b"]) == (77, 77, 77)


def _table_invariants(ex, cfg, n):
    table, excess, blocks = ex["table"], ex["excess"], ex["blocks"]
    nb = cfg.num_buckets
    m = O.OracleMap(cfg)
    seen = set()
    for b in range(nb):
        e0, e1 = table[2 * b], table[2 * b + 1]
        if e0["slot"] < 0:
            assert e1["slot"] < 0 and e1["next"] < 0  # append-only buckets
        chain = [e for e in (e0, e1) if e["slot"] >= 0]
        nx = e1["next"]
        while nx >= 0:
            chain.append(excess[nx])
            assert excess[nx]["slot"] == nx  # overflow entry at its own slot
            nx = excess[nx]["next"]
        for e in chain:
            blk = blocks[e["slot"]]
            assert m.hash(int(blk["x"]), int(blk["y"]), int(blk["z"])) == b
            key = (int(blk["x"]), int(blk["y"]), int(blk["z"]))
            assert key not in seen
            seen.add(key)
    assert len(seen) == n


def test_allocation_set_and_table_invariants():
    c = CFG_SMALL
    cfg = map_cfg(c, num_buckets=1 << 9)   # small table: many overflow chains
    m = O.OracleMap(cfg)
    s = synth(c)
    d, rgb, lab = s.render(0)
    raw, _ = O.filter_depth(d, 1000.0, 3.0)
    m.fuse(raw, rgb, None, None, 0, B.pose_f(np.eye(4)), K(c))
    st = m.stats()
    n = st["allocated_blocks"]
    ex = m.export()
    _table_invariants(ex, cfg, n)
    blocks = ex["blocks"][:n]
    assert np.array_equal(blocks["slot"], np.arange(n))
    keys = set(zip(blocks["x"].tolist(), blocks["y"].tolist(), blocks["z"].tolist()))
    # every valid pixel's surface point lies in an allocated block
    f, cx, cy = c["f"], c["cx"], c["cy"]
    ys, xs = np.nonzero((raw >= 0.1) & (raw <= 3.0))
    z = raw[ys, xs].astype(np.float64)
    px = z * (xs - cx) / f / 0.01
    py = z * (ys - cy) / f / 0.01
    pz = z / 0.01
    for a, b_, cc in zip(np.floor(px).astype(int) // 8, np.floor(py).astype(int) // 8,
                         np.floor(pz).astype(int) // 8):
        assert (a, b_, cc) in keys
    # every allocated block is within T/2 + a block diagonal of some surface point
    pts = np.stack([px, py, pz], 1)
    ctr = np.stack([blocks["x"], blocks["y"], blocks["z"]], 1) * 8 + 4
    from scipy.spatial import cKDTree
    dist, _ = cKDTree(pts).query(ctr)
    assert dist.max() <= 1.5 + 4 * np.sqrt(3) + 1e-6
    # slot 0 holds the block of the first step of the first valid pixel
    y0, x0 = ys[0], xs[0]
    zn = max(0.1, raw[y0, x0] - 0.015)
    first = (int(np.floor(zn * (x0 - cx) / f / 0.01)) // 8,
             int(np.floor(zn * (y0 - cy) / f / 0.01)) // 8, int(np.floor(zn / 0.01)) // 8)
    assert (int(blocks[0]["x"]), int(blocks[0]["y"]), int(blocks[0]["z"])) == first


def test_heap_exhaustion_counted():
    c = CFG_SMALL
    cfg = map_cfg(c, num_blocks=100)
    m = O.OracleMap(cfg)
    d, rgb, lab = synth(c).render(0)
    raw, _ = O.filter_depth(d, 1000.0, 3.0)
    m.fuse(raw, rgb, None, None, 0, B.pose_f(np.eye(4)), K(c))
    st = m.stats()
    assert st["allocated_blocks"] == 100 and st["dropped_heap"] > 0


def test_raycast_wall_depth_and_normal():
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    _wall_fuse(m, 1.0, 0, frames=2)
    k = B.intrinsics(WALL["f"], WALL["f"], WALL["cx"], WALL["cy"])
    zmin, zmax, any_ = m.rendering_blocks(8, 6, 3.0, 0.1, B.pose_f(np.eye(4)), k)
    assert any_
    v, n = m.raycast(64, 48, zmin, zmax, B.pose_f(np.eye(4)), k)
    hit = ~np.isnan(v[..., 0])
    assert hit.mean() > 0.5
    # the half-step refinement (stepScale = 0.5 T / voxel) leaves the hit
    # within a quarter voxel behind the surface
    dz = v[..., 2][hit] - 1.0
    assert dz.min() > -1e-4 and dz.max() < 0.0026
    # the SDF gradient points to the camera
    assert np.allclose(n[..., :3][hit], [0, 0, -1], atol=2e-3)


def test_raycast_empty_map_is_nan():
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    k = B.intrinsics(WALL["f"], WALL["f"], WALL["cx"], WALL["cy"])
    zmin, zmax, any_ = m.rendering_blocks(8, 6, 3.0, 0.1, B.pose_f(np.eye(4)), k)
    assert not any_ and np.all(zmin == 100.0) and np.all(zmax == 0.0)
    v, n = m.raycast(64, 48, zmin, zmax, B.pose_f(np.eye(4)), k)
    assert np.all(np.isnan(v)) and np.all(np.isnan(n))


# -------------------------------------------------- neighbour label pass ---
def _unify_map():
    """A fused wall (allocated, visible blocks), every voxel of every
    allocated block relabelled by the caller."""
    cfg = map_cfg(WALL, num_buckets=WALL["buckets"], num_blocks=WALL["blocks"])
    m = O.OracleMap(cfg)
    _wall_fuse(m, 1.0, 3)
    blocks = m.export()["blocks"][: m.stats()["allocated_blocks"]]
    vis = m.visible()
    assert len(vis) == len(blocks) > 4
    return m, blocks


def _set_all(m, blocks, weight, label, count):
    for b in blocks:
        for i in range(512):
            m.set_label(int(b["slot"]), i, weight, label, count)


def _sem(m, slot, idx):
    s = m.export()["sem"][slot * 512 + idx]
    return int(s["label"]), int(s["count"]), int(s["label2"]), int(s["count2"])


def test_unify_uniform_labels_are_a_fixed_point():
    m, blocks = _unify_map()
    _set_all(m, blocks, 1, 3, 4)
    m.unify_neighbor()
    sem = m.export()["sem"]
    slots = blocks["slot"].astype(np.int64)
    idx = (slots[:, None] * 512 + np.arange(512)[None, :]).ravel()
    assert np.all(sem["label"][idx] == 3) and np.all(sem["count"][idx] == 4)
    assert np.all(sem["label2"][idx] == 0) and np.all(sem["count2"][idx] == 0)


def test_unify_outlier_takes_the_surrounding_label():
    """An interior voxel labelled 7 among voxels labelled 3: it takes 3 with
    the neighbours' largest count and keeps (7, its count) as the secondary
    pair; its neighbours, outvoted 25 to 1, keep 3."""
    m, blocks = _unify_map()
    _set_all(m, blocks, 1, 3, 4)
    slot, idx = int(blocks[0]["slot"]), 4 * 64 + 4 * 8 + 4
    m.set_label(slot, idx, 1, 7, 2)
    m.unify_neighbor()
    assert _sem(m, slot, idx) == (3, 4, 7, 2)
    assert _sem(m, slot, idx + 1) == (3, 4, 0, 0)
    assert _sem(m, slot, idx - 64) == (3, 4, 0, 0)


def _neighbour_hits_self(v, vs):
    """Whether one of the 26 neighbour probes of voxel v lands back on v:
    the probes step in world coordinates, (v * VoxelSize +- VoxelSize) /
    VoxelSize in float, then floor (FuseMap.cu:830-892, DeviceMap.cu:154-163),
    which can round back to v."""
    vs = np.float32(vs)
    p = np.asarray(v, np.float32) * vs
    for d in itertools.product((-1, 0, 1), repeat=3):
        if d == (0, 0, 0):
            continue
        q = p + np.asarray(d, np.float32) * vs
        if np.array_equal(np.floor(q / vs).astype(int), np.asarray(v)):
            return True
    return False


def _isolated_case(m, blocks, want_self):
    b = blocks[0]
    for idx in range(512):
        lx, ly, lz = idx & 7, (idx >> 3) & 7, idx >> 6
        if not (2 <= lx <= 5 and 2 <= ly <= 5 and 2 <= lz <= 5):
            continue
        v = (int(b["x"]) * 8 + lx, int(b["y"]) * 8 + ly, int(b["z"]) * 8 + lz)
        if _neighbour_hits_self(v, WALL["voxel"]) == want_self:
            return int(b["slot"]), idx
    pytest.skip("no such voxel in the block")


def test_unify_isolated_voxel_narrows_minus_one():
    """A labelled voxel with no labelled neighbour: the chosen label is -1,
    stored as 255; its label equals its secondary label, so the count becomes
    count2 + 5 (FuseMap.cu:926-940)."""
    m, blocks = _unify_map()
    _set_all(m, blocks, 0, 0, 0)
    slot, idx = _isolated_case(m, blocks, want_self=False)
    m.set_label(slot, idx, 1, 5, 3, 5, 1)
    m.unify_neighbor()
    assert _sem(m, slot, idx) == (255, 6, 5, 3)


def test_unify_float_probe_can_count_the_voxel_itself():
    """Where a neighbour probe rounds back onto the voxel, the isolated voxel
    counts itself: the chosen label is its own, its count stays."""
    m, blocks = _unify_map()
    _set_all(m, blocks, 0, 0, 0)
    slot, idx = _isolated_case(m, blocks, want_self=True)
    m.set_label(slot, idx, 1, 5, 3, 5, 1)
    m.unify_neighbor()
    assert _sem(m, slot, idx) == (5, 3, 5, 1)


def test_unify_reads_the_state_before_the_pass():
    """Two x-adjacent labelled voxels in different columns, nothing else
    labelled, neither probe landing on itself: each takes the other's label
    as it was before the pass, so they swap (an in-place pass would leave the
    second one unchanged)."""
    m, blocks = _unify_map()
    _set_all(m, blocks, 0, 0, 0)
    b = blocks[0]
    slot = int(b["slot"])
    base = lambda lx, ly, lz: (int(b["x"]) * 8 + lx, int(b["y"]) * 8 + ly, int(b["z"]) * 8 + lz)
    pair = None
    for lz, ly, lx in itertools.product(range(2, 6), range(2, 6), range(2, 5)):
        if not (_neighbour_hits_self(base(lx, ly, lz), WALL["voxel"]) or
                _neighbour_hits_self(base(lx + 1, ly, lz), WALL["voxel"])):
            pair = (lz * 64 + ly * 8 + lx, lz * 64 + ly * 8 + lx + 1)
            break
    assert pair is not None
    a, c = pair
    m.set_label(slot, a, 1, 1, 6)
    m.set_label(slot, c, 1, 2, 9)
    m.unify_neighbor()
    assert _sem(m, slot, a) == (2, 9, 1, 6)
    assert _sem(m, slot, c) == (1, 6, 2, 9)


def test_raycast_pose_derivative_matches_plane_model():
    """Central differences (FP64, h = 1e-6) of the raycast of a fused wall at
    0.9637 m against the closed form for a plane: an exact ray-plane
    intersection moves as n . dv/dxi = [-n, n x v]; the two half-steps
    r += k F(r) dir (RenderScene.cu:465-474) leave the distance to the plane
    times rho = (1 - c/2)^2 (c the ray-normal cosine), so
    n . dv/dxi = (1 - rho) [-n, n x v_hit] - (n . (v_hit - v) / rho) drho/dxi.
    Pins the march's final refinement and its pose dependence."""
    c = dict(cols=64, rows=48, f=60.0, cx=31.5, cy=23.5, voxel=0.01, trunc=0.03,
             buckets=1 << 12, blocks=4000)
    Kc = B.intrinsics(c["f"], c["f"], c["cx"], c["cy"])
    m = O.OracleMap(map_cfg(c))
    rows, cols, z_w = c["rows"], c["cols"], 0.9637
    depth = np.full((rows, cols), z_w, np.float32)
    v = O.compute_vmap(depth, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    n = O.compute_nmap(v)
    n[-1, :] = n[-2, :]
    n[:, -1] = n[:, -2]
    m.fuse(depth, np.full((rows, cols, 3), 100, np.uint8), n, np.ones((rows, cols), np.uint8), 1,
           B.pose_f(np.eye(4)), Kc)
    T = O.se3_exp(np.array([0.003, -0.002, 0.011, 0.0, 0.0, 0.002]))
    zmin, zmax, _ = m.rendering_blocks(8, 6, 3.0, 0.1, B.pose_f(T), Kc)
    v32, _ = m.raycast(cols, rows, zmin, zmax, B.pose_f(T), Kc)
    m64 = m.clone_fp64()
    invf = float(np.float32(1 / c["f"]))
    R, t = T[:3, :3], T[:3, 3]
    nrm = R.T @ np.array([0.0, 0.0, -1.0])
    ez = R.T @ np.array([0.0, 0.0, 1.0])
    h = 1e-6
    Jfd = []
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        vp, _ = m64.raycast_fp64(cols, rows, zmin, zmax, T @ O.se3_exp(e), invf, invf, c["cx"],
                                 c["cy"])
        vm, _ = m64.raycast_fp64(cols, rows, zmin, zmax, T @ O.se3_exp(-e), invf, invf, c["cx"],
                                 c["cy"])
        Jfd.append((vp[..., :3] - vm[..., :3]) @ nrm / (2 * h))
    Jfd = np.stack(Jfd, -1)
    checked = 0
    for y in range(16, 32, 3):
        for x in range(20, 44, 3):
            if np.isnan(v32[y, x, 0]):
                continue
            ray = np.array([(x - c["cx"]) * invf, (y - c["cy"]) * invf, 1.0])
            u = ray / np.linalg.norm(ray)
            cth = -nrm @ u
            rho = (1 - 0.5 * cth) ** 2
            v_hit = u * ((z_w - t[2]) / (ez @ u))
            J = np.concatenate([-nrm, np.cross(nrm, v_hit)])
            drho = np.concatenate([np.zeros(3), -(1 - 0.5 * cth) * np.cross(nrm, u)])
            want = (1 - rho) * J - (nrm @ (v_hit - v32[y, x, :3]) / rho) * drho
            assert np.abs(Jfd[y, x] - want).max() < 1e-4 * np.abs(want).max(), (y, x)
            checked += 1
    assert checked > 30


def test_half_rounding_matches_numpy():
    """The precision form's float -> binary16 rounding (nearest even,
    subnormal results at their own precision, overflow to infinity) against
    numpy's conversion, over normal, subnormal, halfway and special values."""
    rng = np.random.default_rng(17)
    xs = np.concatenate([
        rng.uniform(-1.0, 1.0, 20000),
        rng.uniform(-1e-4, 1e-4, 20000),          # half subnormals
        np.ldexp(rng.integers(1, 1 << 11, 4000) + 0.5, -24 - 10),  # halfway points
        np.float32([0.0, -0.0, 1.0, -1.0, 65504.0, 65520.0, 1e6, -1e6, 6.1e-5, 5.96e-8, 2.98e-8,
                    np.inf, -np.inf]),
    ]).astype(np.float32)
    got = np.array([O.lib().oracle_f32_to_f16(float(x)) for x in xs], np.uint16)
    want = xs.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want), xs[got != want][:10]
    back = np.array([O.lib().oracle_f16_to_f32(int(h)) for h in want], np.float32)
    assert np.array_equal(back, want.view(np.float16).astype(np.float32))
    assert O.lib().oracle_f32_to_f16(float("nan")) & 0x7c00 == 0x7c00
