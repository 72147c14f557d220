"""Parity of the map path (Mapping/DeviceMap.cu, FuseMap.cu, RenderScene.cu)
against the oracle, with the same poses fed to both sides.

Bit-exact: the hash table (both bucket entries and the overflow chains), the
block list, every voxel record (sdf bits, weight, colour, labels, counts),
the visible list, the per-tile depth ranges and the raycast vertex / normal
maps (NaN patterns included).
"""
import numpy as np
import pytest
import torch

from helpers import (B, CFG_SMALL, CFG_VGA, K, bits_equal, empty_dev, host, map_cfg,
                     mismatch_report, synth, to_dev)
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _inputs(c, s, i):
    d, rgb, lab = s.render(i)
    raw, filt = O.filter_depth(d, 1000.0, 3.0)
    v = O.compute_vmap(filt, c["f"], c["f"], c["cx"], c["cy"], 3.0)
    n = O.compute_nmap(v)
    return raw, rgb, n, lab


def _fuse_both(dev, c, frames, gm, om, labels=True, seed=0):
    s = synth(c, seed=seed)
    vis = []
    for i in frames:
        raw, rgb, n, lab = _inputs(c, s, i)
        pose = B.pose_f(s.pose(i))
        nd = 1 if labels else 0
        vh = B.C.c_uint32(0)
        # keep the device tensors referenced until the kernels have run
        tr = [to_dev(a, dev) for a in (raw, rgb, n, lab)]
        B.check(B.lib().sicp_fuse_map_color(gm.h, *[B.image(t) for t in tr], nd, B.C.byref(pose),
                                            None, K(c), B.C.byref(vh), B.stream_ptr()))
        ov = om.fuse(raw, rgb, n, lab, nd, pose, K(c))
        assert vh.value == ov
        vis.append(ov)
    return s, vis


def _compare_maps(gm, om):
    gs, os_ = gm.stats(), om.stats()
    n = os_["allocated_blocks"]
    assert gs["allocated_blocks"] == n
    g = gm.download(n_blocks=n)
    o = om.export()
    assert np.array_equal(g["table"], o["table"]), "hash table differs"
    assert np.array_equal(g["excess"][:n], o["excess"][:n]), "overflow chains differ"
    assert np.array_equal(g["blocks"][:n], o["blocks"][:n]), "block list differs"
    gg, og = g["geom"][: n * 512], o["geom"][: n * 512]
    assert bits_equal(gg["sdf"], og["sdf"]), mismatch_report(gg["sdf"], og["sdf"])
    for f in ("weight", "r", "g", "b"):
        assert np.array_equal(gg[f], og[f]), f
    gsm, osm = g["sem"][: n * 512], o["sem"][: n * 512]
    for f in ("count", "count2", "label", "label2", "r2", "g2", "b2"):
        assert np.array_equal(gsm[f], osm[f]), f
    assert np.array_equal(gm.visible(), om.visible())
    assert gs["dropped_heap"] == os_["dropped_heap"]
    assert gs["dropped_range"] == os_["dropped_range"]
    assert gs["updated_voxels"] == os_["updated_voxels"]
    return n


# The precision form: sdf stored as IEEE half, weights capped at 128.
HALF = dict(tsdf_half=1, max_weight=128)


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("labels", [True, False])
def test_fusion_small_bit_exact(dev, labels, half):
    c = CFG_SMALL
    cfg = map_cfg(c, **(HALF if half else {}))
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, range(0, 12, 2), gm, om, labels=labels)
    n = _compare_maps(gm, om)
    assert n > 500


@pytest.mark.parametrize("half", [False, True])
def test_fusion_vga_bit_exact(dev, half):
    c = CFG_VGA
    cfg = map_cfg(c, **(HALF if half else {}))
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, [0, 3], gm, om)
    n = _compare_maps(gm, om)
    assert n > 2000


@pytest.mark.parametrize("frames", [30, 300])
def test_fusion_vga_sequence_bit_exact(dev, frames):
    """The 640x480 / 5 mm sequence fused frame after frame at its ground-truth
    poses, with labels: after 30 frames (label votes accumulated over every
    frame) and after the whole 300-frame sequence, the table, the overflow
    chains, the block list, both voxel planes with every label and count, and
    the last visible list equal the oracle's."""
    c = CFG_VGA
    cfg = map_cfg(c, num_blocks=60000)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, range(frames), gm, om)
    n = _compare_maps(gm, om)
    assert n > (7000 if frames == 30 else 20000)
    sem = om.export()["sem"][: n * 512]
    assert sem["count"].max() >= min(frames, 255) // 2  # votes accumulated over the frames
    assert (sem["label2"] > 0).sum() > 0


def test_fusion_at_configured_heap_scale_bit_exact(dev):
    """The map at the configured heap size (400k blocks): the 640x480
    sequence fused at 2 mm voxels (every 10th frame of 300, at ground-truth
    poses) allocates about 199k blocks into a 2^16-bucket table (131k
    entries, so about 68k blocks live in overflow chains), with about 100k
    visible blocks per frame. Table, chains, block list, both voxel planes
    and every frame's visible count equal the oracle's."""
    c = dict(CFG_VGA, voxel=0.002, trunc=0.006)
    cfg = map_cfg(c, num_buckets=1 << 16, num_blocks=400000)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _, vis = _fuse_both(dev, c, range(0, 300, 10), gm, om)
    n = _compare_maps(gm, om)
    print("allocated", n, "max visible", max(vis), flush=True)
    assert n > 180000 and max(vis) > 90000
    assert (om.export()["excess"]["slot"] >= 0).sum() > 50000
    assert om.stats()["dropped_heap"] == 0


def test_fusion_overflow_chains_and_heap_exhaustion(dev):
    """A 2^8-bucket table forces long overflow chains; a 700-block heap runs
    out, and both sides must drop the same requests and count them."""
    c = CFG_SMALL
    cfg = map_cfg(c, num_buckets=1 << 8, num_blocks=700)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, [0, 5], gm, om)
    _compare_maps(gm, om)
    assert om.stats()["dropped_heap"] > 0
    assert (om.export()["excess"]["slot"] >= 0).sum() > 100


def test_non_power_of_two_buckets(dev):
    c = CFG_SMALL
    cfg = map_cfg(c, num_buckets=1000003 % 200000 + 100000)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, [0, 4], gm, om)
    _compare_maps(gm, om)


def test_label_count_cap(dev):
    """max_label_count = 255 reproduces the reference's min(255, count + 1)."""
    c = dict(CFG_SMALL, cols=160, rows=120, f=131.25, cx=79.5, cy=59.5)
    cfg = map_cfg(c, max_label_count=255)
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    s = synth(c, noise=False, dropout=0.0)
    raw, rgb, n, lab = _inputs(c, s, 0)
    pose = B.pose_f(np.eye(4))
    tr = [to_dev(a, dev) for a in (raw, rgb, n, lab)]
    for _ in range(300):
        B.check(B.lib().sicp_fuse_map_color(gm.h, *[B.image(t) for t in tr], 1, B.C.byref(pose),
                                            None, K(c), None, B.stream_ptr()))
        om.fuse(raw, rgb, n, lab, 1, pose, K(c))
    _compare_maps(gm, om)
    sem = om.export()["sem"]
    assert sem["count"].max() == 255


def test_empty_frame_no_visible(dev):
    c = CFG_SMALL
    cfg = map_cfg(c)
    gm = B.Map(cfg)
    raw = np.zeros((c["rows"], c["cols"]), np.float32)
    rgb = np.zeros((c["rows"], c["cols"], 3), np.uint8)
    pose = B.pose_f(np.eye(4))
    vh = B.C.c_uint32(7)
    rt, ct = to_dev(raw, dev), to_dev(rgb, dev)
    rc = B.lib().sicp_fuse_map_color(gm.h, B.image(rt), B.image(ct), B.Image(), B.Image(), 0,
                                     B.C.byref(pose), None, K(c), B.C.byref(vh), B.stream_ptr())
    assert rc == B.SICP_ERR_NO_VISIBLE and vh.value == 0


# 1280x720: a 160x90 zRange, above what the prepass gathers in LDS (its
# direct-atomics kernel); the others use the LDS kernel.
CFG_HD = dict(cols=1280, rows=720, f=1050.0, cx=639.5, cy=359.5, voxel=0.008, trunc=0.024,
              buckets=1 << 19, blocks=200000)


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("c", [CFG_SMALL, CFG_VGA, CFG_HD])
def test_prepass_and_raycast_bit_exact(dev, c, half):
    cfg = map_cfg(c, **(HALF if half else {}))
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    s, _ = _fuse_both(dev, c, [0, 2, 4], gm, om)
    pose = B.pose_f(s.pose(4))
    zc, zr = (c["cols"] + 7) // 8, (c["rows"] + 7) // 8
    zmin = empty_dev((zr, zc), torch.float32, dev)
    zmax = empty_dev((zr, zc), torch.float32, dev)
    ne = B.C.c_int(0)
    B.check(B.lib().sicp_create_rendering_blocks(gm.h, B.image(zmin), B.image(zmax), 3.0, 0.1,
                                                 B.C.byref(pose), None, K(c), B.C.byref(ne),
                                                 B.stream_ptr()))
    ozmin, ozmax, oany = om.rendering_blocks(zc, zr, 3.0, 0.1, pose, K(c))
    assert ne.value == 1 and oany
    assert bits_equal(host(zmin), ozmin), mismatch_report(host(zmin), ozmin)
    assert bits_equal(host(zmax), ozmax), mismatch_report(host(zmax), ozmax)
    v = empty_dev((c["rows"], c["cols"], 4), torch.float32, dev, 0, 0.0)
    n = empty_dev((c["rows"], c["cols"], 4), torch.float32, dev, 0, 0.0)
    B.check(B.lib().sicp_raycast(gm.h, B.image(v), B.image(n), B.image(zmin), B.image(zmax),
                                 B.C.byref(pose), None, K(c), B.stream_ptr()))
    ov, on = om.raycast(c["cols"], c["rows"], ozmin, ozmax, pose, K(c))
    assert bits_equal(host(v), ov), mismatch_report(host(v), ov)
    assert bits_equal(host(n), on), mismatch_report(host(n), on)
    hit = ~np.isnan(ov[..., 0])
    assert hit.mean() > 0.3
    if c is CFG_VGA:  # grazing rays reach the second (cooperative) raycast pass
        assert gm.stats()["ray_handover"] > 0


@pytest.mark.parametrize("half", [False, True])
def test_map_snapshot_round_trip(dev, tmp_path, half):
    """sicp_map_save / sicp_map_load (System::WriteMapToDisk / ReadMapFromDisk):
    a map saved after frames 0-4 and loaded into a fresh map continues with
    frames 6-8 exactly as the uninterrupted map does; a snapshot of another
    configuration is refused."""
    c = CFG_SMALL
    cfg = map_cfg(c, **(HALF if half else {}))
    ga, om = B.Map(cfg), O.OracleMap(cfg)
    s, _ = _fuse_both(dev, c, [0, 2, 4], ga, om)
    path = str(tmp_path / "map.bin").encode()
    B.check(B.lib().sicp_map_save(ga.h, path))
    gb = B.Map(cfg)
    B.check(B.lib().sicp_map_load(gb.h, path))
    # every counter but the last raycast's hand-over count (not map state)
    sa, sb = ga.stats(), gb.stats()
    sa.pop("ray_handover")
    sb.pop("ray_handover")
    assert sa == sb
    for m in (ga, gb):
        for i in (6, 8):
            raw, rgb, n, lab = _inputs(c, s, i)
            pose = B.pose_f(s.pose(i))
            tr = [to_dev(a, dev) for a in (raw, rgb, n, lab)]
            B.check(B.lib().sicp_fuse_map_color(m.h, *[B.image(t) for t in tr], 1,
                                                B.C.byref(pose), None, K(c), None,
                                                B.stream_ptr()))
            torch.cuda.synchronize()
    a, b = ga.download(n_blocks=0), gb.download(n_blocks=0)
    na = ga.stats()["allocated_blocks"]
    assert gb.stats()["allocated_blocks"] == na
    for k in ("table", "excess", "blocks"):
        assert np.array_equal(a[k][: len(a[k]) if k == "table" else na],
                              b[k][: len(b[k]) if k == "table" else na]), k
    assert bits_equal(a["geom"]["sdf"][: na * 512], b["geom"]["sdf"][: na * 512])
    assert np.array_equal(a["sem"][: na * 512], b["sem"][: na * 512])
    other = B.Map(map_cfg(dict(c, voxel=0.02)))
    assert B.lib().sicp_map_load(other.h, path) == B.SICP_ERR_INVALID_ARGUMENT


def test_map_load_rejects_corrupt_snapshots(dev, tmp_path):
    """sicp_map_load checks a snapshot on the host before uploading it: a
    truncated file, an entry whose heap slot lies past the used heap, an
    overflow chain that loops and a counter that disagrees with the sections
    are refused with SICP_ERR_INVALID_ARGUMENT, and the map is left reset."""
    c = CFG_SMALL
    cfg = map_cfg(c, num_buckets=1 << 8, num_blocks=4000)  # small table: overflow chains
    gm, om = B.Map(cfg), O.OracleMap(cfg)
    _fuse_both(dev, c, [0, 3], gm, om)
    path = tmp_path / "map.bin"
    B.check(B.lib().sicp_map_save(gm.h, str(path).encode()))
    raw = bytearray(path.read_bytes())
    hdr = int(np.frombuffer(raw[12:16], np.uint32)[0])  # header_bytes
    n = gm.stats()["allocated_blocks"]
    nb2 = 2 * cfg.num_buckets
    table = np.frombuffer(raw[hdr: hdr + 16 * nb2], B.HASH_ENTRY)
    excess_off = hdr + 16 * nb2
    excess = np.frombuffer(raw[excess_off: excess_off + 16 * n], B.HASH_ENTRY)
    # a good copy loads
    good = B.Map(cfg)
    B.check(B.lib().sicp_map_load(good.h, str(path).encode()))
    assert good.stats()["allocated_blocks"] == n

    def refused(data, name):
        p = tmp_path / name
        p.write_bytes(bytes(data))
        m = B.Map(cfg)
        rc = B.lib().sicp_map_load(m.h, str(p).encode())
        st = m.stats()
        assert rc == B.SICP_ERR_INVALID_ARGUMENT, (name, rc)
        assert st["allocated_blocks"] == 0 and st["visible_blocks"] == 0, name
        assert m.download(n_blocks=4, voxels=False)["table"]["slot"].max() == -1, name

    refused(raw[: len(raw) - 100], "truncated.bin")
    refused(raw + b"x", "trailing.bin")
    # an entry pointing past the used heap
    used = np.nonzero(table["slot"] >= 0)[0][0]
    bad = bytearray(raw)
    off = hdr + 16 * used + 8
    bad[off: off + 4] = np.int32(n + 5).tobytes()
    refused(bad, "slot.bin")
    # an overflow chain that loops back on itself
    head = table["next"][1::2]
    chained = np.nonzero(head >= 0)[0]
    assert len(chained) > 0
    e = int(head[chained[0]])
    bad = bytearray(raw)
    off = excess_off + 16 * e + 12
    bad[off: off + 4] = np.int32(e).tobytes()
    refused(bad, "loop.bin")
    # heap_used in the header's counters disagrees with the block section
    bad = bytearray(raw)
    # magic, version, header_bytes, the map config, then the counters
    # (8-byte aligned; heap_used first)
    ctr_off = (16 + B.C.sizeof(B.MapConfig) + 7) // 8 * 8
    bad[ctr_off: ctr_off + 4] = np.uint32(n + 1).tobytes()
    refused(bad, "counter.bin")
    assert excess["slot"][e] == e  # the entry corrupted above was a real chain link
