"""Parity of the pyramid kernels (Tracking/Pyrdown.cu) against the oracle.

Every kernel is compared bit for bit (NaN patterns included) on seeded room
frames, on odd image sizes (border windows of Q7) and on pitched buffers.
"""
import numpy as np
import pytest
import torch

from helpers import (B, CFG_SMALL, CFG_VGA, bits_equal, empty_dev, host, mismatch_report, synth,
                     to_dev)
from oracle import binding as O

pytestmark = pytest.mark.gpu

SIZES = [(CFG_SMALL, 0, 0), (CFG_VGA, 3, 0), (dict(CFG_SMALL, cols=317, rows=239), 0, 5)]


def _frame(c, seed):
    s = synth(c, seed=seed)
    return s.render(seed)


@pytest.mark.parametrize("c,seed,pad", SIZES)
def test_filter_depth(dev, c, seed, pad):
    d, rgb, _ = _frame(c, seed)
    rows, cols = d.shape
    dt = to_dev(d, dev, pad)
    raw = empty_dev((rows, cols), torch.float32, dev, pad)
    filt = empty_dev((rows, cols), torch.float32, dev, pad)
    B.check(B.lib().sicp_filter_depth(B.image(dt), B.image(raw), B.image(filt), 1000.0, 3.0,
                                      B.stream_ptr()))
    oraw, ofilt = O.filter_depth(d, 1000.0, 3.0)
    assert bits_equal(host(raw), oraw), mismatch_report(host(raw), oraw)
    assert bits_equal(host(filt), ofilt), mismatch_report(host(filt), ofilt)


@pytest.mark.parametrize("c,seed,pad", SIZES)
def test_vmap_nmap(dev, c, seed, pad):
    d, _, _ = _frame(c, seed)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    rows, cols = d.shape
    ft = to_dev(filt, dev, pad)
    v = empty_dev((rows, cols, 4), torch.float32, dev, pad)
    n = empty_dev((rows, cols, 4), torch.float32, dev, pad)
    f, cx, cy = c["f"], c["cx"], c["cy"]
    B.check(B.lib().sicp_compute_vmap(B.image(ft), B.image(v), f, f, cx, cy, 3.0, B.stream_ptr()))
    B.check(B.lib().sicp_compute_nmap(B.image(v), B.image(n), B.stream_ptr()))
    ov = O.compute_vmap(filt, f, f, cx, cy, 3.0)
    on = O.compute_nmap(ov)
    assert bits_equal(host(v), ov), mismatch_report(host(v), ov)
    assert bits_equal(host(n), on), mismatch_report(host(n), on)


@pytest.mark.parametrize("c,seed,pad", SIZES)
def test_pyrdown(dev, c, seed, pad):
    d, rgb, _ = _frame(c, seed)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    gray = O.image_to_intensity(rgb)
    rows, cols = d.shape
    src = to_dev(filt, dev, pad)
    dst = empty_dev((rows // 2, cols // 2), torch.float32, dev, pad)
    B.check(B.lib().sicp_pyrdown_gauss_f32(B.image(src), B.image(dst), B.stream_ptr()))
    ref = O.pyrdown_f32(filt)
    assert bits_equal(host(dst), ref), mismatch_report(host(dst), ref)
    g = to_dev(gray, dev, pad)
    gd = empty_dev((rows // 2, cols // 2), torch.uint8, dev, pad)
    B.check(B.lib().sicp_pyrdown_gauss_u8(B.image(g), B.image(gd), B.stream_ptr()))
    assert np.array_equal(host(gd), O.pyrdown_u8(gray))


@pytest.mark.parametrize("c,seed,pad", SIZES)
def test_intensity_derivative(dev, c, seed, pad):
    _, rgb, _ = _frame(c, seed)
    rows, cols = rgb.shape[:2]
    rt = to_dev(rgb, dev, pad)
    g = empty_dev((rows, cols), torch.uint8, dev, pad)
    B.check(B.lib().sicp_image_to_intensity(B.image(rt), B.image(g), B.stream_ptr()))
    og = O.image_to_intensity(rgb)
    assert np.array_equal(host(g), og)
    dx = empty_dev((rows, cols), torch.int16, dev, pad)
    dy = empty_dev((rows, cols), torch.int16, dev, pad)
    B.check(B.lib().sicp_compute_derivative_image(B.image(g), B.image(dx), B.image(dy),
                                                  B.stream_ptr()))
    odx, ody = O.derivative_image(og)
    assert np.array_equal(host(dx), odx)
    assert np.array_equal(host(dy), ody)


@pytest.mark.parametrize("c,seed,pad", SIZES)
def test_resize_map(dev, c, seed, pad):
    d, _, _ = _frame(c, seed)
    _, filt = O.filter_depth(d, 1000.0, 3.0)
    f, cx, cy = c["f"], c["cx"], c["cy"]
    ov = O.compute_vmap(filt, f, f, cx, cy, 3.0)
    on = O.compute_nmap(ov)
    rows, cols = d.shape
    vt, nt = to_dev(ov, dev, pad), to_dev(on, dev, pad)
    vd = empty_dev((rows // 2, cols // 2, 4), torch.float32, dev, pad)
    nd = empty_dev((rows // 2, cols // 2, 4), torch.float32, dev, pad)
    B.check(B.lib().sicp_resize_map(B.image(vt), B.image(nt), B.image(vd), B.image(nd),
                                    B.stream_ptr()))
    rv, rn = O.resize_map(ov, on)
    assert bits_equal(host(vd), rv), mismatch_report(host(vd), rv)
    assert bits_equal(host(nd), rn), mismatch_report(host(nd), rn)


def test_shape_errors(dev):
    a = torch.zeros((10, 10), dtype=torch.float32, device=dev)
    b = torch.zeros((4, 4), dtype=torch.float32, device=dev)
    rc = B.lib().sicp_pyrdown_gauss_f32(B.image(a), B.image(b), B.stream_ptr())
    assert rc == B.SICP_ERR_INVALID_ARGUMENT
