"""Marching-cubes extraction (MeshScene, Mapping/MeshScene.cu:4-356) on the
GPU against the oracle's serial restatement, on maps fused from the same
seeded frames. Bit-exact: the triangle count and, in the canonical order
(heap slot, voxel index, table order), every vertex, normal and colour."""
import numpy as np
import pytest
import torch

from helpers import B, CFG_SMALL, CFG_VGA, bits_equal, map_cfg, mismatch_report
from oracle import binding as O
from test_gpu_map import _fuse_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("c,frames,half", [(CFG_SMALL, [0, 2, 4], 0), (CFG_SMALL, [0, 2, 4], 1),
                                          (CFG_VGA, [0, 3], 0)])
def test_mesh_vs_oracle(dev, c, frames, half):
    cfg = map_cfg(c, tsdf_half=half)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, frames, gm, om)
    tables = B.mesh_tables()
    _, _, _, total = gm.mesh(0)
    ov, on, oc, ototal = om.mesh(tables, total + 16)
    assert total == ototal and total > 1000
    v, n, col, tot2 = gm.mesh(total + 16)
    assert tot2 == total and v.shape[0] == total
    torch.cuda.synchronize()
    v, n, col = v.cpu().numpy(), n.cpu().numpy(), col.cpu().numpy()
    assert bits_equal(v, ov), mismatch_report(v, ov)
    assert bits_equal(n, on), mismatch_report(n, on)
    assert np.array_equal(col, oc)
    # unit normals; vertices inside the allocated blocks' extent
    ln = np.linalg.norm(n.reshape(-1, 3), axis=1)
    assert np.nanmax(np.abs(ln - 1)) < 1e-5
    blocks = gm.download(n_blocks=0, voxels=False)["blocks"]
    na = gm.stats()["allocated_blocks"]
    bw = 8 * c["voxel"]
    lo = np.array([blocks["x"][:na].min(), blocks["y"][:na].min(), blocks["z"][:na].min()]) * bw
    hi = (np.array([blocks["x"][:na].max(), blocks["y"][:na].max(),
                    blocks["z"][:na].max()]) + 1) * bw
    p = v.reshape(-1, 3)
    assert np.all(p >= lo - 1e-6) and np.all(p <= hi + bw)


def test_mesh_cap_and_empty_map(dev):
    """A cap keeps the first triangles of the canonical order and reports the
    number found; an empty map has none."""
    c = CFG_SMALL
    cfg = map_cfg(c)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _, _, _, t0 = gm.mesh(10)
    assert t0 == 0
    _fuse_both(dev, c, [0], gm, om)
    full = gm.mesh(1 << 20)
    part = gm.mesh(777)
    assert part[3] == full[3] and part[0].shape[0] == 777
    torch.cuda.synchronize()
    assert torch.equal(part[0], full[0][:777]) and torch.equal(part[2], full[2][:777])
