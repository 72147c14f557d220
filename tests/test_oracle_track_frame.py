"""The oracle's one-step tracking entry (ComputeSE3 on given pyramids) runs
the same computation as its per-frame system: fed the system's own model and
the frame's measured pyramid, it gives the system's pose to the bit. The
bench's one-step precision figures rest on this entry."""
import numpy as np

from helpers import B, CFG_SMALL, sys_cfg, synth
from oracle import binding as O


def test_track_frame_reproduces_the_system_pose():
    c = CFG_SMALL
    cfg = sys_cfg(c)
    s = synth(c)
    osys = O.OracleSystem(cfg)
    prev_pose = None
    model = None
    checked = 0
    for i in range(5):
        d, rgb, lab = s.render(i)
        r = osys.grab(d, rgb, lab)
        if model is not None:
            nxt = O.measured_pyramid(d, c["f"], c["f"], c["cx"], c["cy"])
            pose, ok = O.track_frame(cfg, nxt, model, prev_pose, fp64=False)
            assert ok == bool(r["tracked"])
            assert np.array_equal(pose, r["pose"]), np.abs(pose - r["pose"]).max()
            # the FP64 transcription lands within tracking precision of it
            p64, ok64 = O.track_frame(cfg, nxt, model, prev_pose, fp64=True)
            assert ok64 and np.abs(p64 - r["pose"]).max() < 1e-3
            checked += 1
        prev_pose = r["pose"]
        model = O.model_pyramid(*osys.model(0))
    assert checked == 4
