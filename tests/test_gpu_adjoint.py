"""The raycast adjoint (sicp_raycast_backward) — the derivative of the
rendered vertex and normal with respect to the camera pose and the TSDF —
against:
  * central finite differences of the oracle's raycast run in FP64 on an
    exact FP64 copy of the same map (zRange fixed; h = 1e-5 at 1 cm voxels,
    1e-7 at 2 mm): the march is piecewise smooth (trilinear cells, discrete
    steps), so a pixel counts when its difference quotients at h and 10 h
    agree, and most pixels must;
  * the analytic point-to-plane (ICP) Jacobian on a fronto-parallel wall: for
    an exact ray-plane intersection n . dv/dxi = [-n, n x v]; the reference's
    two half-steps (r += k F(r) dir twice) stop short by the factor
    rho = (1 - c/2)^2 (c the ray-normal cosine), which gives the model
    n . dv/dxi = (1 - rho) [-n, n x v_hit] - (n . (v_hit - v) / rho) drho/dxi;
  * a torch.autograd.Function wrapper (test harness only).
Tolerances are written in each test."""
import numpy as np
import pytest
import torch

from helpers import B, map_cfg, synth, to_dev
from oracle import binding as O
from test_gpu_map import _inputs

pytestmark = pytest.mark.gpu

ADJ = dict(cols=160, rows=120, f=131.25, cx=79.5, cy=59.5, voxel=0.01, trunc=0.03,
           buckets=1 << 16, blocks=20000)
# 640x480 at the differentiable-renderer workload's 2 mm voxels
ADJ_VGA = dict(cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.002, trunc=0.006,
               buckets=1 << 20, blocks=200000)


def _k(c):
    return B.intrinsics(c["f"], c["f"], c["cx"], c["cy"])


def _invf(c):
    return float(np.float32(1.0 / c["f"]))


def _fuse(dev, c, gm, om, frames_inputs):
    for raw, rgb, n, lab, pose in frames_inputs:
        tr = [to_dev(a, dev) for a in (raw, rgb, n, lab)]
        B.check(B.lib().sicp_fuse_map_color(gm.h, *[B.image(t) for t in tr], 1, B.C.byref(pose),
                                            None, _k(c), None, B.stream_ptr()))
        om.fuse(raw, rgb, n, lab, 1, pose, _k(c))
    torch.cuda.synchronize()


def _scene(dev, c=ADJ):
    cfg = map_cfg(c)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    s = synth(c)
    _fuse(dev, c, gm, om, [(*_inputs(c, s, i), B.pose_f(s.pose(i))) for i in (0, 2, 4)])
    T = s.pose(4)
    zc, zr = (c["cols"] + 7) // 8, (c["rows"] + 7) // 8
    zmin, zmax, _ = om.rendering_blocks(zc, zr, 3.0, 0.1, B.pose_f(T), _k(c))
    return c, gm, om, T, zmin, zmax


class _Backward:
    """One sicp_raycast_backward call site with its device buffers."""

    def __init__(self, dev, c, gm, zmin, zmax, T):
        self.dev, self.c, self.gm = dev, c, gm
        self.zmin, self.zmax = to_dev(zmin, dev), to_dev(zmax, dev)
        self.pose = B.pose_f(T)
        self.scratch = torch.zeros(B.lib().sicp_raycast_backward_scratch_size(),
                                   dtype=torch.uint8, device=dev)

    def __call__(self, pixels, gv, gn=None, want_sdf=False):
        c, dev = self.c, self.dev
        px = torch.as_tensor(np.asarray(pixels, np.int32), device=dev)
        n = px.numel()
        gv_t = torch.as_tensor(np.asarray(gv, np.float32).reshape(n, 3), device=dev)
        gn_t = None if gn is None else torch.as_tensor(np.asarray(gn, np.float32).reshape(n, 3),
                                                       device=dev)
        gpx = torch.zeros((n, 6), dtype=torch.float32, device=dev)
        gpose = torch.zeros(6, dtype=torch.float64, device=dev)
        gsdf = (torch.zeros(self.gm.cfg.num_blocks * 512, dtype=torch.float32, device=dev)
                if want_sdf else None)
        B.check(B.lib().sicp_raycast_backward(
            self.gm.h, B.image(self.zmin), B.image(self.zmax), B.C.byref(self.pose), _invf(c),
            _invf(c), c["cx"], c["cy"], c["cols"], c["rows"], px.data_ptr(), n, gv_t.data_ptr(),
            gn_t.data_ptr() if gn_t is not None else None,
            gsdf.data_ptr() if gsdf is not None else None, gpx.data_ptr(), gpose.data_ptr(),
            self.scratch.data_ptr(), self.scratch.numel(), B.stream_ptr()))
        torch.cuda.synchronize()
        return (gpose.cpu().numpy(), gpx.cpu().numpy(),
                None if gsdf is None else gsdf.cpu().numpy())


def _fd_pose(o64, c, zmin, zmax, T, h):
    """d(vertex, normal)/dxi at every pixel, T' = T exp(xi), FP64."""
    outs = []
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        vp, np_ = o64.raycast_fp64(c["cols"], c["rows"], zmin, zmax, T @ O.se3_exp(e),
                                   _invf(c), _invf(c), c["cx"], c["cy"])
        vm, nm = o64.raycast_fp64(c["cols"], c["rows"], zmin, zmax, T @ O.se3_exp(-e),
                                  _invf(c), _invf(c), c["cx"], c["cy"])
        outs.append(np.concatenate([vp[..., :3] - vm[..., :3], np_[..., :3] - nm[..., :3]], -1)
                    / (2 * h))
    return np.stack(outs, -1)  # rows, cols, 6 (v, n), 6 (xi)


# The difference steps shrink with the voxel: the march is smooth between
# its discrete decisions over a pose range of about a hundredth of a voxel
# at the scene's depth (at 2 mm, steps 1e-5 / 1e-4 agree on 4 % of the hit
# pixels, 1e-7 / 1e-6 on 99 %).
@pytest.mark.parametrize("cam,h", [(ADJ, 1e-5), (ADJ_VGA, 1e-7)], ids=["160x120_1cm", "640x480_2mm"])
def test_pose_jacobian_vs_finite_differences(dev, cam, h):
    c, gm, om, T, zmin, zmax = _scene(dev, cam)
    o64 = om.clone_fp64()
    v32, _ = om.raycast(c["cols"], c["rows"], zmin, zmax, B.pose_f(T), _k(c))
    J1 = _fd_pose(o64, c, zmin, zmax, T, h)
    J2 = _fd_pose(o64, c, zmin, zmax, T, 10 * h)
    hit = ~np.isnan(v32[..., 0])
    smooth = (np.abs(J1 - J2).max((-1, -2)) <= 1e-3 * (np.abs(J1).max((-1, -2)) + 1e-6)) & hit
    assert smooth.sum() > 0.6 * hit.sum()
    rng = np.random.default_rng(3)
    ys, xs = np.nonzero(hit)
    pick = rng.choice(len(ys), 400, replace=False)
    pix = ys[pick] * c["cols"] + xs[pick]
    bw = _Backward(dev, c, gm, zmin, zmax, T)
    gpu = np.zeros((len(pix), 6, 6))
    for j in range(6):  # rows of d(v, n)/dxi through unit upstream gradients
        g = np.zeros((len(pix), 3), np.float32)
        g[:, j % 3] = 1
        if j < 3:
            gpose, gpx, _ = bw(pix, g)
        else:
            gpose, gpx, _ = bw(pix, np.zeros_like(g), g)
        gpu[:, j, :] = gpx
        assert np.allclose(gpose, gpx.astype(np.float64).sum(0), rtol=1e-5, atol=1e-6)
    fd = J1[ys[pick], xs[pick]]
    sm = smooth[ys[pick], xs[pick]]
    err = np.abs(gpu - fd).max((-1, -2)) / (np.abs(fd).max((-1, -2)) + 1e-6)
    # FP32 analytic derivative against FP64 differences, where those are smooth
    assert np.mean(err[sm] < 1e-2) > 0.97, np.sort(err[sm])[-20:]
    assert np.median(err[sm]) < 1e-3


def test_sdf_gradient_vs_finite_differences(dev):
    c, gm, om, T, zmin, zmax = _scene(dev)
    o64 = om.clone_fp64()
    v32, _ = om.raycast(c["cols"], c["rows"], zmin, zmax, B.pose_f(T), _k(c))
    blocks = om.export()["blocks"]
    na = om.stats()["allocated_blocks"]
    assert np.array_equal(gm.download(n_blocks=0, voxels=False)["blocks"][:na], blocks[:na])
    bw = _Backward(dev, c, gm, zmin, zmax, T)
    ys, xs = np.nonzero(~np.isnan(v32[..., 0]))
    rng = np.random.default_rng(5)
    checked = 0
    for i in rng.choice(len(ys), 12, replace=False):
        y, x = ys[i], xs[i]
        _, _, gsdf = bw([y * c["cols"] + x], [[0, 0, 1]], want_sdf=True)
        idx = np.nonzero(gsdf)[0]
        if len(idx) == 0:
            continue
        fd = np.zeros(len(idx))
        for q, k in enumerate(idx):
            slot, li = int(k // 512), int(k % 512)
            s0 = o64.get_sdf(slot, li)
            vals = []
            for h in (1e-5, -1e-5):
                o64.set_sdf(slot, li, s0 + h)
                v, _ = o64.raycast_fp64(c["cols"], c["rows"], zmin, zmax, T, _invf(c), _invf(c),
                                        c["cx"], c["cy"])
                vals.append(v[y, x, 2])
            o64.set_sdf(slot, li, s0)
            fd[q] = (vals[0] - vals[1]) / 2e-5
        scale = np.abs(fd).max() + 1e-9
        if np.isnan(fd).any():
            continue
        assert np.abs(gsdf[idx] - fd).max() / scale < 1e-2, (y, x, gsdf[idx], fd)
        checked += 1
    assert checked >= 8


def _wall(dev, z_w, T):
    """A wall at depth z_w fused from the identity pose, its zRange at T."""
    c = dict(cols=64, rows=48, f=60.0, cx=31.5, cy=23.5, voxel=0.01, trunc=0.03,
             buckets=1 << 12, blocks=4000)
    cfg = map_cfg(c)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    rows, cols = c["rows"], c["cols"]
    depth = np.full((rows, cols), z_w, np.float32)
    v = O.compute_vmap(depth, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    n = O.compute_nmap(v)
    n[-1, :] = n[-2, :]
    n[:, -1] = n[:, -2]
    rgb = np.full((rows, cols, 3), 100, np.uint8)
    lab = np.ones((rows, cols), np.uint8)
    _fuse(dev, c, gm, om, [(depth, rgb, n, lab, B.pose_f(np.eye(4)))])
    zc, zr = (cols + 7) // 8, (rows + 7) // 8
    zmin, zmax, _ = om.rendering_blocks(zc, zr, 3.0, 0.1, B.pose_f(T), _k(c))
    return c, gm, om, zmin, zmax


def test_wall_matches_the_analytic_icp_jacobian(dev):
    """Point-to-plane derivative of the rendered vertex of a wall at 0.9637 m
    (8.4 voxels behind the front face of its nearest allocated block, so the
    march's block-width steps do not change where it meets the surface),
    seen from a camera moved off the voxel lattice, against the closed form
    above with the wall's normal n = R^T (0, 0, -1) in the camera frame."""
    z_w = 0.9637
    T = O.se3_exp(np.array([0.003, -0.002, 0.011, 0.0, 0.0, 0.002]))
    c, gm, om, zmin, zmax = _wall(dev, z_w, T)
    v32, _ = om.raycast(c["cols"], c["rows"], zmin, zmax, B.pose_f(T), _k(c))
    R, t = T[:3, :3], T[:3, 3]
    nrm = R.T @ np.array([0.0, 0.0, -1.0])
    ez = R.T @ np.array([0.0, 0.0, 1.0])
    pix, model = [], []
    for y in range(16, 32, 3):
        for x in range(20, 44, 3):
            if np.isnan(v32[y, x, 0]):
                continue
            ray = np.array([(x - c["cx"]) * _invf(c), (y - c["cy"]) * _invf(c), 1.0])
            u = ray / np.linalg.norm(ray)
            cth = -nrm @ u
            rho = (1 - 0.5 * cth) ** 2
            v_hit = u * ((z_w - t[2]) / (ez @ u))
            J = np.concatenate([-nrm, np.cross(nrm, v_hit)])
            drho = np.concatenate([np.zeros(3), -(1 - 0.5 * cth) * np.cross(nrm, u)])
            m = (1 - rho) * J - (nrm @ (v_hit - v32[y, x, :3]) / rho) * drho
            pix.append(y * c["cols"] + x)
            model.append(m)
    assert len(pix) > 30
    bw = _Backward(dev, c, gm, zmin, zmax, T)
    _, gpx, _ = bw(pix, np.tile(nrm.astype(np.float32), (len(pix), 1)))
    model = np.array(model)
    err = np.abs(gpx - model).max(1) / np.abs(model).max(1)
    assert err.max() < 2e-3, err
    # along the view axis: three quarters of the camera's approach
    assert abs(model[:, 2].max() - 0.75) < 0.01


class _Render(torch.autograd.Function):
    """Sampled vertices and normals of the raycast as a function of a pose
    twist (applied as T exp(xi); evaluated at xi = 0) and of the TSDF, whose
    values the map holds (the `sdf` argument stands for them)."""

    @staticmethod
    def forward(ctx, xi, sdf, bw, vfull, nfull, pix):
        ctx.bw, ctx.pix = bw, pix
        idx = torch.as_tensor(pix, device=vfull.device, dtype=torch.long)
        return vfull.reshape(-1, 4)[idx, :3].clone(), nfull.reshape(-1, 4)[idx, :3].clone()

    @staticmethod
    def backward(ctx, gv, gn):
        gpose, _, gsdf = ctx.bw(ctx.pix, gv.detach().cpu().numpy(), gn.detach().cpu().numpy(),
                                want_sdf=True)
        return (torch.as_tensor(gpose), torch.as_tensor(gsdf, device=gv.device), None, None, None,
                None)


def test_autograd_wrapper(dev):
    c, gm, om, T, zmin, zmax = _scene(dev)
    vfull = torch.empty((c["rows"], c["cols"], 4), dtype=torch.float32, device=dev)
    nfull = torch.empty_like(vfull)
    zt = [to_dev(a, dev) for a in (zmin, zmax)]
    pose = B.pose_f(T)
    B.check(B.lib().sicp_raycast_inv(gm.h, B.image(vfull), B.image(nfull), B.image(zt[0]),
                                     B.image(zt[1]), B.C.byref(pose), None, _invf(c), _invf(c),
                                     c["cx"], c["cy"], B.stream_ptr()))
    torch.cuda.synchronize()
    v_h = vfull.cpu().numpy()
    ys, xs = np.nonzero(~np.isnan(v_h[..., 0]))
    rng = np.random.default_rng(9)
    sel = rng.choice(len(ys), 300, replace=False)
    pix = ys[sel] * c["cols"] + xs[sel]
    bw = _Backward(dev, c, gm, zmin, zmax, T)
    xi = torch.zeros(6, dtype=torch.float64, requires_grad=True)
    sdf = torch.zeros(gm.cfg.num_blocks * 512, dtype=torch.float32, device=dev, requires_grad=True)
    v, n = _Render.apply(xi, sdf, bw, vfull, nfull, pix)
    w = torch.as_tensor(np.random.default_rng(1).normal(size=(len(pix), 3)), dtype=torch.float32,
                        device=dev)
    loss = (w * v).sum() + (0.3 * w * n).sum()
    loss.backward()
    # the same contraction from the per-pixel Jacobians through the ABI directly
    gpose, _, gsdf = bw(pix, w.cpu().numpy(), 0.3 * w.cpu().numpy(), want_sdf=True)
    assert np.allclose(xi.grad.numpy(), gpose, rtol=1e-6, atol=1e-9)
    assert torch.count_nonzero(sdf.grad) > 0
    assert np.allclose(sdf.grad.cpu().numpy(), gsdf, rtol=1e-4, atol=1e-6)
