"""sicp_system_stage / sicp_batch_stage: the frame copied into the staged
inputs by one kernel equals its source byte for byte, whatever the row
alignment (16-byte rows, 4-byte rows, odd byte counts, padded pitches); shape
mismatches are refused before any copy."""
import numpy as np
import pytest
import torch

from helpers import B, CFG_SMALL, download, sys_cfg, to_dev

pytestmark = pytest.mark.gpu


def _cfg(cols, rows):
    return sys_cfg(dict(CFG_SMALL, cols=cols, rows=rows, cx=(cols - 1) / 2.0,
                        cy=(rows - 1) / 2.0))


def _frame(cols, rows, seed, pad=0):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, 65535, (rows, cols), dtype=np.uint16)
    c = rng.integers(0, 255, (rows, cols, 3), dtype=np.uint8)
    lab = rng.integers(0, 10, (rows, cols), dtype=np.uint8)
    dev = torch.device("cuda:0")
    return (d, c, lab), tuple(to_dev(a, dev, pad) for a in (d, c, lab))


def _staged(g, cols, rows):
    d, c, lab = g.staging()
    return (download(d, rows, cols, np.uint16), download(c, rows, cols, np.uint8, 3),
            download(lab, rows, cols, np.uint8))


@pytest.mark.parametrize("cols,rows,pad", [(640, 480, 0), (322, 241, 0), (321, 17, 0),
                                           (320, 240, 3)])
def test_stage_copies_bytes(dev, cols, rows, pad):
    """640: every row 16-byte aligned; 322: depth rows of 644 bytes (4-byte
    units), rgb rows of 966 (bytes); 321: odd everywhere; pad: pitched
    sources (pitch = 323 elements)."""
    print("stage", cols, rows, pad, flush=True)
    g = B.System(_cfg(cols, rows))
    host, devs = _frame(cols, rows, cols + rows, pad)
    g.stage(*devs)
    torch.cuda.synchronize()
    for got, want in zip(_staged(g, cols, rows), host):
        assert np.array_equal(got, want)
    # without labels: the label image keeps its previous contents
    host2, devs2 = _frame(cols, rows, 7, pad)
    g.stage(devs2[0], devs2[1], None)
    torch.cuda.synchronize()
    got = _staged(g, cols, rows)
    assert np.array_equal(got[0], host2[0]) and np.array_equal(got[1], host2[1])
    assert np.array_equal(got[2], host[2])
    g.close()


def test_batch_stage_copies_every_system(dev):
    sizes = [(640, 480), (320, 240), (322, 241)]
    systems = [B.System(_cfg(c, r)) for c, r in sizes]
    b = B.Batch(systems)
    frames = [_frame(c, r, 100 + q) for q, (c, r) in enumerate(sizes)]
    n = len(sizes)
    dep, rgb, lab = (B.Image * n)(), (B.Image * n)(), (B.Image * n)()
    for q, (_, (d, c, l)) in enumerate(frames):
        dep[q], rgb[q] = B.image(d), B.image(c)
        lab[q] = B.image(l if q != 1 else None)  # system 1: no labels
    B.check(B.lib().sicp_batch_stage(b.h, dep, rgb, lab, B.stream_ptr()))
    torch.cuda.synchronize()
    for q, ((c, r), (host, _)) in enumerate(zip(sizes, frames)):
        got = _staged(systems[q], c, r)
        assert np.array_equal(got[0], host[0]) and np.array_equal(got[1], host[1]), q
        if q != 1:
            assert np.array_equal(got[2], host[2]), q
    b.close()
    for g in systems:
        g.close()


def test_stage_refuses_shape_mismatch(dev):
    g = B.System(_cfg(320, 240))
    _, (d, c, lab) = _frame(320, 240, 1)
    _, (d2, c2, lab2) = _frame(321, 240, 2)
    L = B.lib()
    I = lambda t: B.C.byref(B.image(t))  # noqa: E731
    for args in ((I(d2), I(c), I(lab)), (I(d), I(c2), I(lab)), (I(d), I(c), I(lab2)),
                 (None, I(c), None), (I(d), None, None)):
        assert L.sicp_system_stage(g.h, *args, None) == B.SICP_ERR_INVALID_ARGUMENT
    assert L.sicp_system_stage(None, I(d), I(c), None, None) == B.SICP_ERR_INVALID_ARGUMENT
    g.close()
