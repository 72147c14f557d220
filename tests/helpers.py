"""Shared helpers of the parity tests: seeded frames, device upload, NaN-aware
bit comparisons."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from semantic_icpslam_amd import binding as B  # noqa: E402

# Camera / map configurations (the small CPU case and the 640x480 case)
CFG_SMALL = dict(cols=320, rows=240, f=262.5, cx=159.5, cy=119.5, voxel=0.01, trunc=0.03,
                 buckets=1 << 18, blocks=60000)
CFG_VGA = dict(cols=640, rows=480, f=525.0, cx=319.5, cy=239.5, voxel=0.005, trunc=0.015,
               buckets=1 << 20, blocks=400000)


def synth(c, seed=0, noise=True, dropout=0.02):
    return B.Synth(c["cols"], c["rows"], c["f"], c["f"], c["cx"], c["cy"], seed=seed,
                   noise=noise, dropout=dropout)


def map_cfg(c, **kw):
    cfg = B.map_config(c["voxel"], c["trunc"], c["buckets"], c["blocks"])
    cfg.max_cols, cfg.max_rows = c["cols"], c["rows"]
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def sys_cfg(c, **kw):
    cfg = B.system_config(c["cols"], c["rows"], c["f"], c["f"], c["cx"], c["cy"], c["voxel"],
                          c["trunc"], c["buckets"], c["blocks"])
    for k, v in kw.items():
        setattr(cfg.map, k, v)
    return cfg


def K(c, level=0):
    s = float(1 << level)
    return B.intrinsics(c["f"] / s, c["f"] / s, c["cx"] / s, c["cy"] / s)


def to_dev(a, dev, pad_cols=0):
    """Uploads a (rows, cols[, ch]) array; pad_cols > 0 gives a pitched view."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if pad_cols:
        shape = list(t.shape)
        shape[1] += pad_cols
        big = torch.zeros(shape, dtype=t.dtype)
        big[:, : a.shape[1]] = t
        return big.to(dev)[:, : a.shape[1]]
    return t.to(dev)


def empty_dev(shape, dtype, dev, pad_cols=0, fill=None):
    import torch
    shape = list(shape)
    cols = shape[1]
    shape[1] += pad_cols
    t = torch.empty(shape, dtype=dtype, device=dev)
    if fill is not None:
        t.fill_(fill)
    return t[:, :cols] if pad_cols else t


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bits_equal(a, b):
    """Bitwise equality of float arrays, with every NaN equal to every NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        na, nb = np.isnan(a), np.isnan(b)
        if not np.array_equal(na, nb):
            return False
        ai = a.view(np.uint32 if a.dtype == np.float32 else np.uint64)
        bi = b.view(np.uint32 if b.dtype == np.float32 else np.uint64)
        return bool(np.array_equal(ai[~na], bi[~nb]))
    return bool(np.array_equal(a, b))


def mismatch_report(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    both = ~na & ~nb
    d = np.abs(a[both] - b[both])
    return (f"nan-pattern diff {int((na != nb).sum())}, value diff {int((d > 0).sum())}, "
            f"max {d.max() if d.size else 0}")


def download(im, rows, cols, dtype, channels=1):
    """Copies a pitched device image (a binding.Image) into a numpy array."""
    import ctypes
    import torch
    torch.cuda.synchronize()
    dt = np.dtype(dtype)
    shape = (rows, cols, channels) if channels > 1 else (rows, cols)
    buf = np.zeros(shape, dt)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    width = cols * channels * dt.itemsize
    assert hip.hipMemcpy2D(buf.ctypes.data, width, im.ptr, im.pitch, width, rows, 2) == 0
    return buf
