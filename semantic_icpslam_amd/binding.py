"""ctypes binding of the C ABI in include/sicp.h (libsicp.so).

The library is the product: every function here calls the compiled HIP code.
Loading fails loudly when libsicp.so is missing; there is no fallback path.
Device buffers are torch tensors on a ROCm device (torch is used for memory
and streams only); `image()` turns a tensor into the pitched `sicp_image`
descriptor the ABI takes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SICP_LIB names another build of the same sources (the profiling build).
LIB_PATH = os.environ.get("SICP_LIB") or os.path.join(_HERE, "libsicp.so")

SICP_OK = 0
SICP_ERR_INVALID_ARGUMENT = 1
SICP_ERR_HIP = 2
SICP_ERR_CAPACITY = 3
SICP_ERR_NO_VISIBLE = 4
SICP_ERR_NOTHING_TO_RENDER = 5


class Image(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("pitch", C.c_size_t), ("cols", C.c_int), ("rows", C.c_int)]


class Intrinsics(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float)]


class PoseF(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("Rinv", C.c_float * 9), ("t", C.c_float * 3)]


class MapConfig(C.Structure):
    _fields_ = [
        ("voxel_size", C.c_float), ("truncation", C.c_float),
        ("depth_min", C.c_float), ("depth_max", C.c_float),
        ("num_buckets", C.c_int32), ("num_blocks", C.c_int32),
        ("max_weight", C.c_int32), ("max_label_count", C.c_int32),
        ("request_capacity", C.c_int32), ("max_cols", C.c_int32), ("max_rows", C.c_int32),
        ("unify_labels", C.c_int32), ("tsdf_half", C.c_int32),
    ]


class MapStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "allocated_blocks", "visible_blocks", "new_blocks", "dropped_heap",
        "dropped_requests", "dropped_range", "updated_voxels", "render_tiles", "ray_handover")] + [
        (n, C.c_uint64) for n in ("updated_total", "visible_total", "fusions",
                                  "heap_scanned_total")]


class Tuning(C.Structure):
    """sicp_tuning: the raycast's launch shape."""
    _fields_ = [("ray_first_pass_samples", C.c_int32), ("ray_second_pass_workgroups", C.c_int32),
                ("ray_lds_block_table", C.c_int32)]


class SystemConfig(C.Structure):
    _fields_ = [
        ("cols", C.c_int32), ("rows", C.c_int32), ("K", Intrinsics),
        ("depth_scale", C.c_float), ("depth_cutoff", C.c_float),
        ("icp_iters", C.c_int32 * 3), ("icp_min_count", C.c_int32 * 3),
        ("icp_thresh", C.c_float), ("map", MapConfig),
        ("icp_only", C.c_int32), ("rgb_weight", C.c_float), ("icp_sums_fp32", C.c_int32),
        ("icp_spread", C.c_int32),
    ]


class FrameResult(C.Structure):
    _fields_ = [
        ("frame_index", C.c_int32), ("tracked", C.c_int32), ("icp_count", C.c_int32),
        ("iterations", C.c_int32), ("icp_error", C.c_float), ("visible_blocks", C.c_uint32),
        ("allocated_blocks", C.c_uint32), ("pose", C.c_double * 16),
        ("rgb_error", C.c_float), ("rgb_count", C.c_int32),
    ]


class SynthConfig(C.Structure):
    _fields_ = [
        ("cols", C.c_int32), ("rows", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
        ("cx", C.c_float), ("cy", C.c_float), ("seed", C.c_uint64),
        ("depth_units_per_m", C.c_float), ("dropout", C.c_float), ("noise", C.c_int32),
    ]


# numpy mirrors of the map records
HASH_ENTRY = np.dtype([("key", "<u8"), ("slot", "<i4"), ("next", "<i4")])
BLOCK = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4"), ("slot", "<i4")])
VOXEL_GEOM = np.dtype([("sdf", "<f4"), ("weight", "u1"), ("r", "u1"), ("g", "u1"), ("b", "u1")])
VOXEL_SEM = np.dtype([("count", "<u2"), ("count2", "<u2"), ("label", "u1"), ("label2", "u1"),
                      ("r2", "u1"), ("g2", "u1"), ("b2", "u1"), ("pad", "u1", 3)])
assert HASH_ENTRY.itemsize == 16 and BLOCK.itemsize == 16
assert VOXEL_GEOM.itemsize == 8 and VOXEL_SEM.itemsize == 12

_lib = None

_P = C.c_void_p
_SIGS = {
    "sicp_last_error": (C.c_char_p, []),
    "sicp_version": (C.c_char_p, []),
    "sicp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "sicp_filter_depth": (C.c_int, [Image, Image, Image, C.c_float, C.c_float, _P]),
    "sicp_compute_vmap": (C.c_int, [Image, Image, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, _P]),
    "sicp_compute_nmap": (C.c_int, [Image, Image, _P]),
    "sicp_pyrdown_gauss_f32": (C.c_int, [Image, Image, _P]),
    "sicp_pyrdown_gauss_u8": (C.c_int, [Image, Image, _P]),
    "sicp_image_to_intensity": (C.c_int, [Image, Image, _P]),
    "sicp_compute_derivative_image": (C.c_int, [Image, Image, Image, _P]),
    "sicp_resize_map": (C.c_int, [Image, Image, Image, Image, _P]),
    "sicp_icp_step": (C.c_int, [Image, Image, Image, Image, C.POINTER(PoseF), C.POINTER(PoseF),
                                Intrinsics, _P, _P, _P, _P, _P, _P, _P]),
    "sicp_icp_step_fp32": (C.c_int, [Image, Image, Image, Image, C.POINTER(PoseF),
                                     C.POINTER(PoseF), Intrinsics, _P, _P, _P, _P, _P, _P, _P]),
    "sicp_rgb_step": (C.c_int, [Image, Image, Image, Image, Image, Image, C.POINTER(PoseF),
                                C.POINTER(PoseF), Intrinsics, _P, C.c_size_t, _P, _P, _P, _P,
                                _P, _P, _P, _P]),
    "sicp_rgb_scratch_size": (C.c_size_t, [C.c_int32, C.c_int32]),
    "sicp_so3_step": (C.c_int, [Image, Image, Image, Image, C.POINTER(PoseF), C.POINTER(PoseF),
                                Intrinsics, _P, _P, _P, _P, _P, _P, _P]),
    "sicp_mesh_scene": (C.c_int, [_P, _P, _P, _P, C.c_uint32, _P, _P, _P]),
    "sicp_unify_neighbor_labels": (C.c_int, [_P, _P]),
    "sicp_raycast_backward": (C.c_int, [_P, Image, Image, C.POINTER(PoseF), C.c_float, C.c_float,
                                        C.c_float, C.c_float, C.c_int32, C.c_int32, _P,
                                        C.c_int32, _P, _P, _P, _P, _P, _P, C.c_size_t, _P]),
    "sicp_raycast_backward_scratch_size": (C.c_size_t, []),
    "sicp_mesh_tables": (C.c_int, [_P, _P, _P]),
    "sicp_mesh_write_ply": (C.c_int, [C.c_char_p, _P, _P, _P, C.c_uint32]),
    "sicp_icp_partials_size": (C.c_size_t, []),
    "sicp_icp_step_static": (C.c_int, [Image, Image, Image, Image, C.POINTER(PoseF),
                                       C.POINTER(PoseF), Intrinsics, _P, _P, _P, _P, _P, _P]),
    "sicp_se3_update_host": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "sicp_map_config_default": (None, [C.POINTER(MapConfig), C.c_float, C.c_float, C.c_int32,
                                       C.c_int32]),
    "sicp_map_create": (C.c_int, [C.POINTER(MapConfig), C.POINTER(_P)]),
    "sicp_map_destroy": (C.c_int, [_P]),
    "sicp_map_get_config": (C.c_int, [_P, C.POINTER(MapConfig)]),
    "sicp_map_reset": (C.c_int, [_P, _P]),
    "sicp_fuse_map_color": (C.c_int, [_P, Image, Image, Image, Image, C.c_int, C.POINTER(PoseF),
                                      _P, Intrinsics, C.POINTER(C.c_uint32), _P]),
    "sicp_check_block_visibility": (C.c_int, [_P, C.POINTER(PoseF), _P, C.c_int, C.c_int,
                                              Intrinsics, C.c_float, C.c_float,
                                              C.POINTER(C.c_uint32), _P]),
    "sicp_create_rendering_blocks": (C.c_int, [_P, Image, Image, C.c_float, C.c_float,
                                               C.POINTER(PoseF), _P, Intrinsics,
                                               C.POINTER(C.c_int), _P]),
    "sicp_raycast": (C.c_int, [_P, Image, Image, Image, Image, C.POINTER(PoseF), _P, Intrinsics,
                               _P]),
    "sicp_raycast_inv": (C.c_int, [_P, Image, Image, Image, Image, C.POINTER(PoseF), _P,
                                   C.c_float, C.c_float, C.c_float, C.c_float, _P]),
    "sicp_map_visible_list": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "sicp_map_get_stats": (C.c_int, [_P, C.POINTER(MapStats), _P]),
    "sicp_map_save": (C.c_int, [_P, C.c_char_p]),
    "sicp_map_load": (C.c_int, [_P, C.c_char_p]),
    "sicp_map_download": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_uint32]),
    "sicp_map_download_visible": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "sicp_system_config_default": (None, [C.POINTER(SystemConfig), C.c_int, C.c_int, C.c_float,
                                          C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_int32, C.c_int32]),
    "sicp_system_create": (C.c_int, [C.POINTER(SystemConfig), C.POINTER(_P)]),
    "sicp_system_destroy": (C.c_int, [_P]),
    "sicp_system_grab_image": (C.c_int, [_P, Image, Image, Image, C.c_int, _P]),
    "sicp_system_last_result": (C.c_int, [_P, C.POINTER(FrameResult), _P]),
    "sicp_system_map": (_P, [_P]),
    "sicp_system_set_timing": (C.c_int, [_P, C.c_int]),
    "sicp_system_read_timing": (C.c_int, [_P, _P, _P, C.c_int, _P]),
    "sicp_system_staging": (C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image)]),
    "sicp_system_grab_staged": (C.c_int, [_P, C.c_int, C.c_int, _P]),
    "sicp_system_stage": (C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image),
                                    _P]),
    "sicp_system_read_timing_detail": (C.c_int, [_P, C.c_int, _P, C.c_int, C.POINTER(C.c_int),
                                                 _P]),
    "sicp_system_frame_image": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Image)]),
    "sicp_system_grab_batch": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P]),
    "sicp_batch_create": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "sicp_batch_destroy": (C.c_int, [_P]),
    "sicp_batch_grab_staged": (C.c_int, [_P, _P, C.c_int, _P]),
    "sicp_batch_stage": (C.c_int, [_P, C.POINTER(Image), C.POINTER(Image), C.POINTER(Image), _P]),
    "sicp_batch_set_shared_tracking": (C.c_int, [_P, C.c_int]),
    "sicp_system_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "sicp_system_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "sicp_system_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(Tuning), C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), _P]),
    "sicp_system_autotune_reset": (C.c_int, [_P]),
    "sicp_system_autotune_results": (C.c_int, [_P, _P, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "sicp_synth_create": (C.c_int, [C.POINTER(SynthConfig), C.POINTER(_P)]),
    "sicp_synth_destroy": (C.c_int, [_P]),
    "sicp_synth_pose": (C.c_int, [_P, C.c_int, _P]),
    "sicp_synth_render": (C.c_int, [_P, C.c_int, _P, _P, _P]),
}

EXPORTED = tuple(_SIGS.keys())



def mesh_tables():
    """The marching-cubes tables sicp_mesh_scene uses: edge (256,),
    triangle (256, 16), vertex count (256,). Host only."""
    edge = np.zeros(256, np.int32)
    tri = np.zeros((256, 16), np.int8)
    nvert = np.zeros(256, np.int32)
    check(lib().sicp_mesh_tables(edge.ctypes.data, tri.ctypes.data, nvert.ctypes.data),
          "sicp_mesh_tables")
    return edge, tri, nvert


def lib():
    """Loads libsicp.so (once). Raises if the library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                "g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SicpError(RuntimeError):
    def __init__(self, code, what):
        msg = lib().sicp_last_error().decode(errors="replace")
        super().__init__(f"{what} failed with status {code}: {msg}")
        self.code = code


def check(code, what="sicp call"):
    if code != SICP_OK:
        raise SicpError(code, what)
    return code


# ------------------------------------------------------------ descriptors --
def image(t, cols=None, rows=None):
    """sicp_image of a (rows, cols[, channels]) tensor; rows may be padded
    (pitch = stride of dim 0)."""
    if t is None:
        return Image(None, 0, 0, 0)
    assert t.stride(-1) == 1 or t.dim() == 2
    if t.dim() == 3:
        assert t.stride(1) == t.shape[2] and t.stride(2) == 1, "pixels must be contiguous"
    else:
        assert t.stride(1) == 1
    pitch = t.stride(0) * t.element_size()
    return Image(t.data_ptr(), pitch, cols if cols is not None else t.shape[1],
                 rows if rows is not None else t.shape[0])


def intrinsics(fx, fy, cx, cy):
    return Intrinsics(fx, fy, cx, cy)


def pose_f(T):
    """Float view of a 4x4 camera->world pose: R, R^T, t (Frame::GpuRotation)."""
    T = np.asarray(T, dtype=np.float64)
    p = PoseF()
    R = T[:3, :3]
    for i in range(3):
        for j in range(3):
            p.R[i * 3 + j] = float(np.float32(R[i, j]))
            p.Rinv[i * 3 + j] = float(np.float32(R[j, i]))
        p.t[i] = float(np.float32(T[i, 3]))
    return p


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def device_count():
    n = C.c_int(0)
    lib().sicp_device_count(C.byref(n))
    return n.value


# ---------------------------------------------------------------- map API --
def map_config(voxel_size, truncation, num_buckets, num_blocks, **kw):
    cfg = MapConfig()
    lib().sicp_map_config_default(C.byref(cfg), voxel_size, truncation, num_buckets, num_blocks)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def map_stats(handle, stream=None):
    """Counters of a map handle (owned by a Map or by a System)."""
    s = MapStats()
    check(lib().sicp_map_get_stats(handle, C.byref(s), stream_ptr(stream)), "map_get_stats")
    return {n: getattr(s, n) for n, _ in MapStats._fields_}


class Map:
    """Owner of one device map (the buffers Mapping::Create allocates)."""

    def __init__(self, cfg: MapConfig):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().sicp_map_create(C.byref(cfg), C.byref(h)), "sicp_map_create")
        self.h = h

    def close(self):
        if self.h:
            lib().sicp_map_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None):
        check(lib().sicp_map_reset(self.h, stream_ptr(stream)), "sicp_map_reset")

    def stats(self, stream=None):
        return map_stats(self.h, stream)

    def mesh(self, max_triangles, stream=None):
        """MeshScene into device tensors: (vertices, normals, colours) of
        shape (n, 3, 3) on the current device, and the number found."""
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        v = torch.empty((max(max_triangles, 1), 3, 3), dtype=torch.float32, device=dev)
        n = torch.empty_like(v)
        c = torch.empty((max(max_triangles, 1), 3, 3), dtype=torch.uint8, device=dev)
        k, tot = C.c_uint32(0), C.c_uint32(0)
        check(lib().sicp_mesh_scene(self.h, v.data_ptr(), n.data_ptr(), c.data_ptr(),
                                    max_triangles, C.byref(k), C.byref(tot), stream_ptr(stream)),
              "sicp_mesh_scene")
        return v[: k.value], n[: k.value], c[: k.value], tot.value

    def download(self, n_blocks=0, voxels=True):
        cfg = self.cfg
        nblk = cfg.num_blocks if not n_blocks else min(n_blocks, cfg.num_blocks)
        table = np.zeros(2 * cfg.num_buckets, HASH_ENTRY)
        excess = np.zeros(nblk, HASH_ENTRY)
        blocks = np.zeros(nblk, BLOCK)
        geom = np.zeros(nblk * 512, VOXEL_GEOM) if voxels else None
        sem = np.zeros(nblk * 512, VOXEL_SEM) if voxels else None
        ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a is not None else None
        check(lib().sicp_map_download(self.h, ptr(table), ptr(excess), ptr(blocks), ptr(geom),
                                      ptr(sem), nblk), "sicp_map_download")
        return dict(table=table, excess=excess, blocks=blocks, geom=geom, sem=sem)

    def visible(self):
        n = C.c_uint32(0)
        check(lib().sicp_map_download_visible(self.h, None, 0, C.byref(n)), "download_visible")
        out = np.zeros(n.value, BLOCK)
        if n.value:
            check(lib().sicp_map_download_visible(self.h, out.ctypes.data_as(C.c_void_p), n.value,
                                                  C.byref(n)), "download_visible")
        return out


# ------------------------------------------------------------- synthetic --
class Synth:
    """Seeded procedural room (test / benchmark input)."""

    def __init__(self, cols, rows, fx, fy, cx, cy, seed=0, noise=True, dropout=0.02,
                 depth_units_per_m=1000.0):
        cfg = SynthConfig(cols, rows, fx, fy, cx, cy, seed, depth_units_per_m, dropout,
                          1 if noise else 0)
        h = C.c_void_p()
        check(lib().sicp_synth_create(C.byref(cfg), C.byref(h)), "sicp_synth_create")
        self.h, self.cols, self.rows = h, cols, rows

    def __del__(self):
        try:
            if self.h:
                lib().sicp_synth_destroy(self.h)
        except Exception:
            pass

    def pose(self, i):
        T = np.zeros(16, np.float64)
        check(lib().sicp_synth_pose(self.h, i, T.ctypes.data_as(C.c_void_p)), "synth_pose")
        return T.reshape(4, 4)

    def render(self, i):
        d = np.zeros((self.rows, self.cols), np.uint16)
        rgb = np.zeros((self.rows, self.cols, 3), np.uint8)
        lab = np.zeros((self.rows, self.cols), np.uint8)
        check(lib().sicp_synth_render(self.h, i, d.ctypes.data_as(C.c_void_p),
                                      rgb.ctypes.data_as(C.c_void_p),
                                      lab.ctypes.data_as(C.c_void_p)), "synth_render")
        return d, rgb, lab


# ----------------------------------------------------------------- system --
def system_config(cols, rows, fx, fy, cx, cy, voxel_size, truncation, num_buckets, num_blocks,
                  **map_kw):
    cfg = SystemConfig()
    lib().sicp_system_config_default(C.byref(cfg), cols, rows, fx, fy, cx, cy, voxel_size,
                                     truncation, num_buckets, num_blocks)
    for k, v in map_kw.items():
        setattr(cfg.map, k, v)
    return cfg


class System:
    """The per-frame driver (System::GrabImage counterpart)."""

    def __init__(self, cfg: SystemConfig):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().sicp_system_create(C.byref(cfg), C.byref(h)), "sicp_system_create")
        self.h = h

    def close(self):
        if self.h:
            lib().sicp_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grab(self, depth, rgb, labels=None, stream=None):
        nd = 1 if labels is not None else 0
        check(lib().sicp_system_grab_image(self.h, image(depth), image(rgb), image(labels), nd,
                                           stream_ptr(stream)), "sicp_system_grab_image")

    def staging(self):
        """The system's staged input images (depth, rgb, labels) for grab_staged."""
        ims = [Image(), Image(), Image()]
        check(lib().sicp_system_staging(self.h, *[C.byref(i) for i in ims]), "system_staging")
        return ims

    def stage(self, depth, rgb, labels=None, stream=None):
        """Copies the frame into the staged inputs on `stream` (one kernel)."""
        lab = image(labels) if labels is not None else None
        check(lib().sicp_system_stage(self.h, C.byref(image(depth)), C.byref(image(rgb)),
                                      C.byref(lab) if lab is not None else None,
                                      stream_ptr(stream)), "system_stage")

    def grab_staged(self, depth, rgb, labels=None, use_graph=True, stream=None):
        """Copies the frame into the staged inputs on `stream` (device to device)
        and runs GrabImage from them; from the second frame on as one graph."""
        self.stage(depth, rgb, labels, stream)
        nd = 1 if labels is not None else 0
        check(lib().sicp_system_grab_staged(self.h, nd, 1 if use_graph else 0, stream_ptr(stream)),
              "system_grab_staged")

    def result(self, stream=None):
        r = FrameResult()
        check(lib().sicp_system_last_result(self.h, C.byref(r), stream_ptr(stream)),
              "system_last_result")
        return dict(frame_index=r.frame_index, tracked=r.tracked, icp_count=r.icp_count,
                    iterations=r.iterations, icp_error=r.icp_error,
                    visible_blocks=r.visible_blocks, allocated_blocks=r.allocated_blocks,
                    pose=np.array(list(r.pose)).reshape(4, 4),
                    rgb_error=r.rgb_error, rgb_count=r.rgb_count)

    def map_handle(self):
        return lib().sicp_system_map(self.h)

    def frame_image(self, which, level=0):
        im = Image()
        check(lib().sicp_system_frame_image(self.h, which, level, C.byref(im)), "frame_image")
        return im

    def tuning(self):
        t = Tuning()
        check(lib().sicp_system_get_tuning(self.h, C.byref(t)), "system_get_tuning")
        return dict(ray_first_pass_samples=t.ray_first_pass_samples,
                    ray_second_pass_workgroups=t.ray_second_pass_workgroups,
                    ray_lds_block_table=t.ray_lds_block_table)

    def set_tuning(self, ray_first_pass_samples, ray_second_pass_workgroups,
                   ray_lds_block_table=1):
        t = Tuning(ray_first_pass_samples, ray_second_pass_workgroups, ray_lds_block_table)
        check(lib().sicp_system_set_tuning(self.h, C.byref(t)), "system_set_tuning")

    def autotune(self, repeats=5, stream=None):
        """Times the last frame's raycast under each candidate launch shape,
        adds the medians to the system's sums and keeps the lowest sum
        (sicp_system_autotune); returns the choice with its mean time per
        call and the default shape's, in microseconds."""
        t = Tuning()
        us, dflt = C.c_float(), C.c_float()
        check(lib().sicp_system_autotune(self.h, repeats, C.byref(t), C.byref(us), C.byref(dflt),
                                         stream_ptr(stream)), "system_autotune")
        return dict(ray_first_pass_samples=t.ray_first_pass_samples,
                    ray_second_pass_workgroups=t.ray_second_pass_workgroups,
                    ray_lds_block_table=t.ray_lds_block_table,
                    raycast_us=round(us.value, 2), default_us=round(dflt.value, 2))

    def autotune_results(self):
        """Every autotuner candidate with its mean raycast time (microseconds)."""
        shapes, us, n = (Tuning * 64)(), (C.c_float * 64)(), C.c_int32()
        check(lib().sicp_system_autotune_results(self.h, shapes, us, 64, C.byref(n)),
              "system_autotune_results")
        return [(shapes[i].ray_first_pass_samples, shapes[i].ray_second_pass_workgroups,
                 round(us[i], 2)) for i in range(n.value)]


def grab_batch(systems, frames, stream=None):
    """One frame of every system in shared tracking launches
    (sicp_system_grab_batch); `frames` holds one (depth, rgb, labels or None)
    tuple of device tensors per system."""
    n = len(systems)
    hs = (C.c_void_p * n)(*[g.h for g in systems])
    depth = (Image * n)(*[image(f[0]) for f in frames])
    rgb = (Image * n)(*[image(f[1]) for f in frames])
    lab = (Image * n)(*[image(f[2]) for f in frames])
    nd = (C.c_int32 * n)(*[1 if f[2] is not None else 0 for f in frames])
    check(lib().sicp_system_grab_batch(hs, n, depth, rgb, lab, nd, stream_ptr(stream)),
          "sicp_system_grab_batch")


class Batch:
    """Systems advanced together, each round (one frame of every system) as
    one hipGraph launch once every system is past its first frame."""

    def __init__(self, systems):
        self.systems = list(systems)
        n = len(self.systems)
        h = C.c_void_p()
        check(lib().sicp_batch_create((C.c_void_p * n)(*[g.h for g in self.systems]), n,
                                      C.byref(h)), "sicp_batch_create")
        self.h = h

    def close(self):
        if self.h:
            lib().sicp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shared_tracking(self, shared):
        """True: the systems' tracking iterations share launches; False: each
        system's frame is an independent chain of the round."""
        check(lib().sicp_batch_set_shared_tracking(self.h, 1 if shared else 0),
              "sicp_batch_set_shared_tracking")

    def grab_staged(self, frames, use_graph=True, stream=None):
        """Copies every system's frame into its staged inputs on `stream` and
        runs the round."""
        n = len(self.systems)
        dep, rgb, lab = (Image * n)(), (Image * n)(), (Image * n)()
        for q, f in enumerate(frames):
            dep[q], rgb[q], lab[q] = image(f[0]), image(f[1]), image(f[2])
        check(lib().sicp_batch_stage(self.h, dep, rgb, lab, stream_ptr(stream)), "batch_stage")
        nd = (C.c_int32 * n)(*[1 if f[2] is not None else 0 for f in frames])
        check(lib().sicp_batch_grab_staged(self.h, nd, 1 if use_graph else 0,
                                           stream_ptr(stream)), "sicp_batch_grab_staged")
