"""Pins of the oracle's ComputeSE3 decisions (Tracking/Tracking.cc:398-516)
that do not come from its own code: error streams whose outcomes were traced
by hand through the reference's loop, and the final motion gate against
scipy's Euler decomposition.

The loop runs the levels 2, 1, 0 with 3, 5 and 10 iterations (iteration
indices 0-2, 3-7, 8-17) and the correspondence minimums 100, 1000, 2000.
The trace replaces the pose by a marker: 0 is the frame's initial pose and
k + 1 the pose after iteration k's update.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import binding as O

OK = 5000  # a count above every level's minimum


def _stream(errors, counts=None):
    e = np.asarray(errors, np.float32)
    c = np.full(len(e), OK, np.int32) if counts is None else np.asarray(counts, np.int32)
    return e, c


def test_all_iterations_applied_when_the_error_falls():
    e, c = _stream(np.linspace(1.0, 0.1, 18))
    marker, failed, applied, ran, err = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (18, 0, 18, 18)
    assert err == np.float32(0.1)


def test_count_below_the_level_minimum_fails_and_rolls_back():
    # iteration 3 is the first of level 1, whose minimum is 1000: 999 fails
    e, c = _stream(np.linspace(1.0, 0.5, 18), [OK, OK, OK, 999] + [OK] * 14)
    marker, failed, applied, ran, _ = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (0, 1, 3, 4)
    # the same count passes at level 2 (minimum 100)
    e, c = _stream(np.linspace(1.0, 0.5, 18), [999] * 3 + [OK] * 15)
    assert O.se3_control_trace(e, c)[:4] == (18, 0, 18, 18)
    # exactly the minimum passes (the test is count < minimum)
    e, c = _stream(np.linspace(1.0, 0.5, 18), [100, 100, 100, 1000] + [OK] * 4 + [2000] * 10)
    assert O.se3_control_trace(e, c)[:4] == (18, 0, 18, 18)


def test_nan_error_fails():
    e, c = _stream([1.0, 0.9, np.nan] + [0.5] * 15)
    marker, failed, applied, ran, _ = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (0, 1, 2, 3)


def test_three_increases_restore_the_candidate_with_the_last_increase_error():
    # k=0 1.0 applied (pose 1); k=1 0.9 applied (pose 2);
    # k=2 0.95 > 0.9: first increase, candidate = pose 2 (the pose whose
    #     error rose), candidateICP = 0.95, applied (pose 3);
    # k=3 0.97 > 0.95: second increase, candidateICP = 0.97 (no braces),
    #     applied (pose 4);
    # k=4 0.99 > 0.97: third increase -> pose = candidate 2, error = 0.99.
    e, c = _stream([1.0, 0.9, 0.95, 0.97, 0.99] + [0.1] * 13)
    marker, failed, applied, ran, err = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (2, 0, 4, 5)
    # with braces around both statements the error would be 0.95
    assert err == np.float32(0.99)


def test_a_fall_between_increases_restarts_the_count():
    # k=1 up (candidate pose 1), k=2 down (count 0), k=3 up (candidate pose 3),
    # k=4 up, k=5 up -> restore pose 3 with the error of k=5
    e, c = _stream([1.0, 1.1, 1.05, 1.2, 1.3, 1.4] + [0.1] * 12)
    marker, failed, applied, ran, err = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (3, 0, 5, 6)
    assert err == np.float32(1.4)


def test_the_increase_count_and_last_error_carry_across_levels():
    # level 2: k=0 0.5, k=1 0.6 (up, candidate pose 1), k=2 0.7 (up);
    # level 1: k=3 0.8 (third increase) -> break out of every level
    e, c = _stream([0.5, 0.6, 0.7, 0.8] + [0.1] * 14)
    marker, failed, applied, ran, err = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (1, 0, 3, 4)
    assert err == np.float32(0.8)


def test_first_error_above_the_initial_100_counts_as_an_increase():
    # lastError starts at 100 (Tracking.cc:391): 150 is an increase, the
    # candidate is the initial pose
    e, c = _stream([150.0, 160.0, 170.0] + [0.1] * 15)
    marker, failed, applied, ran, err = O.se3_control_trace(e, c)
    assert (marker, failed, applied, ran) == (0, 0, 2, 3)
    assert err == np.float32(170.0)


def test_equal_error_is_not_an_increase():
    e, c = _stream([0.5] * 18)
    assert O.se3_control_trace(e, c)[:4] == (18, 0, 18, 18)


def test_photometric_term_enters_the_error_and_fails_first():
    # currError = rgbError * w + icpError: a constant ICP error with a rising
    # photometric error counts as an increase
    e, c = _stream([0.5] * 18)
    rgb = np.array([100.0, 200.0, 300.0, 400.0] + [0.0] * 14, np.float32)
    rc = np.full(18, OK, np.int32)
    marker, failed, applied, ran, err = O.se3_control_trace(e, c, rgb_error=rgb, rgb_count=rc,
                                                            w=1e-3)
    assert (marker, failed, applied, ran) == (1, 0, 3, 4)
    # the photometric count is tested before ICP's (Tracking.cc:425-432)
    rc[2] = 10
    marker, failed, applied, ran, _ = O.se3_control_trace(e, c, rgb_error=rgb, rgb_count=rc,
                                                          w=1e-3)
    assert (marker, failed, applied, ran) == (0, 1, 2, 3)


# ------------------------------------------------------------ final gate --
def _pose(rotvec, t):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(rotvec).as_matrix()
    T[:3, 3] = t
    return T


def _sine_norm_scipy(T):
    """|sin| of the intrinsic X-Y-Z Euler angles (Eigen's eulerAngles(0, 1, 2),
    R = Rx(a) Ry(b) Rz(c)); both solutions give the same norm."""
    a = Rotation.from_matrix(T[:3, :3]).as_euler("XYZ")
    return float(np.linalg.norm(np.sin(a)))


@pytest.mark.parametrize("seed", range(6))
def test_euler_sine_norm_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    T0 = _pose(rng.normal(size=3) * 0.5, rng.normal(size=3))
    T1 = T0 @ _pose(rng.normal(size=3) * 0.2, rng.normal(size=3) * 0.1)
    rn, tn = O.euler_sin_norm(T0, T1)
    rel = np.linalg.inv(T0) @ T1
    assert abs(rn - _sine_norm_scipy(rel)) < 1e-12
    assert abs(tn - np.linalg.norm(rel[:3, 3])) < 1e-12


def test_gate_thresholds():
    T0 = _pose([0.1, -0.2, 0.3], [1.0, 2.0, 0.5])
    small = T0 @ _pose([0.02, 0.01, -0.03], [0.03, -0.02, 0.05])
    assert O.se3_gate(5e-5, 1e-4, T0, small)
    # error not strictly below the threshold
    assert not O.se3_gate(1e-4, 1e-4, T0, small)
    # translation 0.12 m > 0.1
    assert not O.se3_gate(5e-5, 1e-4, T0, T0 @ _pose([0, 0, 0], [0.0, 0.12, 0.0]))
    # translation 0.099 m passes
    assert O.se3_gate(5e-5, 1e-4, T0, T0 @ _pose([0, 0, 0], [0.0, 0.0, 0.099]))
    # rotation of 6 degrees about x: sin = 0.1045 > 0.1
    assert not O.se3_gate(5e-5, 1e-4, T0, T0 @ _pose([np.radians(6.0), 0, 0], [0, 0, 0]))
    # 5.7 degrees: sin = 0.0993
    assert O.se3_gate(5e-5, 1e-4, T0, T0 @ _pose([np.radians(5.7), 0, 0], [0, 0, 0]))
    # two axes at 4.5 degrees: norm of sines 0.111 > 0.1 although each is below
    two = Rotation.from_euler("XYZ", [np.radians(4.5), np.radians(4.5), 0.0]).as_matrix()
    T = np.eye(4)
    T[:3, :3] = two
    assert not O.se3_gate(5e-5, 1e-4, T0, T0 @ T)


# ------------------------------------------------ forced motion-gate failure --
def test_pose_jump_fails_the_motion_gate():
    """Frame 0 then frame 9 of the synthetic trajectory (true motion 0.104 m):
    tracking converges (ICP error below the threshold, all 18 iterations) yet
    the frame is lost and the pose restored, which only the translation gate
    (|t| > 0.1 m, Tracking.cc:505-516) can do; frame 8 (0.093 m) tracks."""
    from helpers import CFG_SMALL, sys_cfg, synth
    c = CFG_SMALL
    cfg = sys_cfg(c)
    s = synth(c)
    for jump, tracked in ((9, 0), (8, 1)):
        motion = np.linalg.norm((np.linalg.inv(s.pose(0)) @ s.pose(jump))[:3, 3])
        assert (motion > 0.1) == (tracked == 0)
        o = O.OracleSystem(cfg)
        r0 = o.grab(*s.render(0))
        r = o.grab(*s.render(jump))
        assert r["tracked"] == tracked, (jump, r)
        assert 0 < r["icp_error"] < cfg.icp_thresh and r["icp_count"] > 2000
        if tracked:
            est = np.linalg.norm((np.linalg.inv(r0["pose"]) @ r["pose"])[:3, 3])
            assert abs(est - motion) < 0.01 and r["visible_blocks"] > 0
        else:
            assert r["iterations"] == 18
            assert np.array_equal(r["pose"], r0["pose"])
            assert r["visible_blocks"] == 0 and r["allocated_blocks"] == r0["allocated_blocks"]
