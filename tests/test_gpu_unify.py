"""The neighbour label pass (UnifyNeighbor, Mapping/FuseMap.cu:757-948) on
the GPU against the oracle: after fusion with the pass (the reference's
FuseMapColor sequence) and through the stand-alone entry point, the label
planes are bit-exact. The pass is deterministic: every voxel reads the state
before the pass (a snapshot of the blocks being rewritten)."""
import numpy as np
import pytest

from helpers import B, CFG_SMALL, CFG_VGA, map_cfg, sys_cfg, synth, to_dev
from oracle import binding as O
from test_gpu_map import _compare_maps, _fuse_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("c,frames", [(CFG_SMALL, range(0, 10, 2)), (CFG_VGA, [0, 2, 4])])
def test_fusion_with_unify_bit_exact(dev, c, frames):
    cfg = map_cfg(c, unify_labels=1, max_label_count=255)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, frames, gm, om)
    _compare_maps(gm, om)
    # the pass did change labels: compare with a map fused without it
    ref = O.OracleMap(map_cfg(c, max_label_count=255))
    s = synth(c)
    from test_gpu_map import _inputs
    from helpers import K
    for i in frames:
        raw, rgb, n, lab = _inputs(c, s, i)
        ref.fuse(raw, rgb, n, lab, 1, B.pose_f(s.pose(i)), K(c))
    na = om.stats()["allocated_blocks"]
    a, b = om.export()["sem"][: na * 512], ref.export()["sem"][: na * 512]
    assert np.count_nonzero(a["label"] != b["label"]) > 0


def test_unify_entry_point_twice(dev):
    """sicp_unify_neighbor_labels on a map fused without the pass, twice in a
    row (the second pass starts from the first one's result)."""
    c = CFG_SMALL
    cfg = map_cfg(c)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, [0, 3, 6], gm, om)
    for _ in range(2):
        B.check(B.lib().sicp_unify_neighbor_labels(gm.h, B.stream_ptr()))
        om.unify_neighbor()
        _compare_maps(gm, om)


def test_system_with_unify_matches_oracle(dev):
    c = CFG_SMALL
    cfg = sys_cfg(c, unify_labels=1)
    g, o = B.System(cfg), O.OracleSystem(cfg)
    s = synth(c)
    for i in range(5):
        d, rgb, lab = s.render(i)
        g.grab(to_dev(d, dev), to_dev(rgb, dev), to_dev(lab, dev))
        rg, ro = g.result(), o.grab(d, rgb, lab)
        assert rg["tracked"] == ro["tracked"]
        assert np.abs(rg["pose"] - ro["pose"]).max() < 1e-4
        assert abs(rg["allocated_blocks"] - ro["allocated_blocks"]) <= 0.01 * ro["allocated_blocks"] + 2
