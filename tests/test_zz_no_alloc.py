"""sicp_icp_step_static (ICPStep with the partial rows in library-held device
memory, no allocation) against sicp_icp_step (the same kernels with a
caller workspace), through the C ABI: the 29 sums and the expanded A, b and
residual are bitwise equal on two calls in a row, the second on a smaller
image, and again with the calls in the other order."""
import numpy as np
import pytest
import torch

from helpers import B, CFG_VGA, synth
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _maps(dev, frame, level):
    c = CFG_VGA
    s = synth(c)
    d, _, _ = s.render(frame)
    v, n = O.measured_pyramid(d, c["f"], c["f"], c["cx"], c["cy"])[level]
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (v, n)], s


def _call(fn, args, partials=None):
    sums = torch.zeros(29, dtype=torch.float32, device=args[0].device)
    host = np.zeros(29, np.float32)
    A = np.zeros(36, np.float64)
    b = np.zeros(6, np.float64)
    res = np.zeros(2, np.float32)
    vc, vl, nc, nl, pc, pl, K = args
    head = [B.image(vc), B.image(vl), B.image(nc), B.image(nl), B.C.byref(pc), B.C.byref(pl), K]
    tail = [sums.data_ptr(), host.ctypes.data, A.ctypes.data, b.ctypes.data, res.ctypes.data,
            B.stream_ptr()]
    if partials is not None:
        B.check(fn(*head, partials.data_ptr(), *tail), "sicp_icp_step")
    else:
        B.check(fn(*head, *tail), "sicp_icp_step_static")
    torch.cuda.synchronize()
    return sums.cpu().numpy(), host, A, b, res


def _case(dev, level, static_first):
    (vc, nc), s = _maps(dev, 3, level)
    (vl, nl), _ = _maps(dev, 1, level)
    c = CFG_VGA
    f = 2.0 ** level
    K = B.intrinsics(c["f"] / f, c["f"] / f, c["cx"] / f, c["cy"] / f)
    pc, pl = B.pose_f(s.pose(2)), B.pose_f(s.pose(1))
    args = (vc, vl, nc, nl, pc, pl, K)
    partials = torch.zeros(B.lib().sicp_icp_partials_size(), dtype=torch.uint8, device=dev)
    lib = B.lib()
    if static_first:
        a = _call(lib.sicp_icp_step_static, args)
        b = _call(lib.sicp_icp_step, args, partials)
    else:
        b = _call(lib.sicp_icp_step, args, partials)
        a = _call(lib.sicp_icp_step_static, args)
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert a[1][28] > 1000  # correspondences found
    return a


@pytest.mark.parametrize("static_first", [True, False])
def test_static_scratch_equals_workspace_entry(dev, static_first):
    big = _case(dev, 0, static_first)    # 640 x 480
    small = _case(dev, 1, static_first)  # then 320 x 240: the same array, fewer rows
    assert big[1][28] > small[1][28]


def test_static_entry_rejects_oversize_images(dev):
    im = B.Image(B.C.c_void_p(256), 16 * 40000, 40000, 40000)  # 1.6e9 pixels
    p = B.pose_f(np.eye(4))
    K = B.intrinsics(1.0, 1.0, 0.0, 0.0)
    rc = B.lib().sicp_icp_step_static(im, im, im, im, B.C.byref(p), B.C.byref(p), K, None, None,
                                      None, None, None, None)
    assert rc == B.SICP_ERR_INVALID_ARGUMENT
    assert b"2^30" in B.lib().sicp_last_error()
