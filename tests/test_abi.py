"""The C ABI library loads and exports every entry point include/sicp.h
declares; argument errors come back as status codes (no GPU needed: these
calls are rejected before any HIP call)."""
import ctypes as C
import os
import re
import subprocess

from helpers import B, ROOT

HEADER = os.path.join(ROOT, "include", "sicp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(sicp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = B.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sicp_\w+)", out))
    assert set(names) <= exported


def test_binding_covers_header():
    assert set(B.EXPORTED) <= set(declared_functions())


def test_invalid_arguments_return_status():
    lib = B.lib()
    cfg = B.map_config(0.005, 0.015, 1 << 10, 100)
    cfg.request_capacity = 1000  # not a power of two
    h = C.c_void_p()
    assert lib.sicp_map_create(C.byref(cfg), C.byref(h)) == B.SICP_ERR_INVALID_ARGUMENT
    assert b"power of two" in lib.sicp_last_error()
    cfg = B.map_config(-1.0, 0.015, 1 << 10, 100)
    assert lib.sicp_map_create(C.byref(cfg), C.byref(h)) == B.SICP_ERR_INVALID_ARGUMENT
    a = B.Image(C.c_void_p(16), 40, 10, 10)
    b = B.Image(C.c_void_p(16), 16, 4, 4)
    assert lib.sicp_pyrdown_gauss_f32(a, b, None) == B.SICP_ERR_INVALID_ARGUMENT
    assert lib.sicp_filter_depth(a, a, b, 1000.0, 3.0, None) == B.SICP_ERR_INVALID_ARGUMENT
    bad_pitch = B.Image(C.c_void_p(16), 8, 10, 10)  # pitch < cols * 4
    assert lib.sicp_compute_nmap(bad_pitch, bad_pitch, None) == B.SICP_ERR_INVALID_ARGUMENT
    sc = B.SystemConfig()
    lib.sicp_system_config_default(C.byref(sc), 4, 4, 1.0, 1.0, 2.0, 2.0, 0.01, 0.03, 64, 10)
    assert lib.sicp_system_create(C.byref(sc), C.byref(h)) == B.SICP_ERR_INVALID_ARGUMENT


def test_config_defaults_follow_reference_constants():
    cfg = B.map_config(0.004, 0.024, 1000000, 50000)
    # DeviceMap.h:168-169 depth range; FuseMap.cu:505 weight cap
    assert abs(cfg.depth_min - 0.1) < 1e-7 and abs(cfg.depth_max - 3.0) < 1e-7
    assert cfg.max_weight == 255
    sc = B.system_config(640, 480, 525.0, 525.0, 319.5, 239.5, 0.005, 0.015, 1 << 20, 1000)
    # Tracking.h:110,112,117
    assert list(sc.icp_iters) == [10, 5, 3]
    assert list(sc.icp_min_count) == [2000, 1000, 100]
    assert abs(sc.icp_thresh - 1e-4) < 1e-10


def test_synthetic_room_is_deterministic():
    s1 = B.Synth(80, 60, 65.6, 65.6, 39.5, 29.5, seed=3)
    s2 = B.Synth(80, 60, 65.6, 65.6, 39.5, 29.5, seed=3)
    a, b = s1.render(7), s2.render(7)
    for x, y in zip(a, b):
        assert (x == y).all()
    d, rgb, lab = a
    assert (d > 0).mean() > 0.9 and lab.max() >= 1
    assert abs(s1.pose(0) - __import__("numpy").eye(4)).max() == 0


def test_reference_named_layer_builds_and_loads():
    """The reference-named C++ layer compiles against the C ABI alone (the
    test driver links only libsicp.so and the HIP runtime) and loads without
    a GPU."""
    import ctypes
    lib = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libmirror_check.so")
    assert os.path.exists(lib), "run build() first"
    h = ctypes.CDLL(lib)
    for name in ("mirror_frame", "mirror_icp", "mirror_map"):
        assert hasattr(h, name), name
