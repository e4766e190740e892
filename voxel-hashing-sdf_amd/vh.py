"""ctypes binding of libvh (include/vh.h, include/vh_synth.h).

Python-side mirror of the engine's host interface for tests, benchmarks and scripts. The
reference's host API is the C++ class ark::GpuTsdfGenerator (include/tsdf.cuh:548-683);
`Volume` keeps its three operations - construct, integrate(depth, pose), extract mesh -
and adds the query / backward / Gauss-Newton extensions.

Loading order matters: PyTorch wheels carry their own HIP runtime (SONAME
libamdhip64.so.7). Importing torch before libvh makes both resolve to that one runtime, so
device pointers from torch tensors are valid inside libvh. `load()` does this and refuses to
run with two HIP runtimes mapped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VH_LIBRARY selects another build of the library (for comparing builds); default: in-tree
LIB_PATH = os.environ.get("VH_LIBRARY") or os.path.join(_HERE, "libvh.so")

VH_OK = 0
VH_ERR_INVALID_ARG = 1
VH_ERR_OOM = 2
VH_ERR_POOL_FULL = 3
VH_ERR_HASH_FULL = 4
VH_ERR_DEVICE = 5
VH_ERR_RANGE = 6
VH_ERR_TRUNCATED = 7
VH_ERR_IO = 8
VH_ERR_NO_DEVICE = 9
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "POOL_FULL", 4: "HASH_FULL", 5: "DEVICE",
                6: "RANGE", 7: "TRUNCATED", 8: "IO", 9: "NO_DEVICE"}

STORAGE_F32 = 0
STORAGE_F16 = 1
STORAGE_F16R = 2
DEPTH_U16 = 0
DEPTH_F32 = 1
PRIOR_ABSOLUTE = 0
PRIOR_INCREMENT = 1
ABI_VERSION = 7
# per-evaluation solver stats: n_valid, rms, |delta xi|, lambda, accepted, accepted pose (12)
GN_STATS = 17


class VhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libvh status {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float), ("fy", C.c_float),
                ("cx", C.c_float), ("cy", C.c_float), ("max_depth", C.c_float), ("min_depth", C.c_float),
                ("vox_size", C.c_float), ("trunc", C.c_float), ("log2_buckets", C.c_int32),
                ("pool_blocks", C.c_int32), ("storage", C.c_int32), ("weight_cap", C.c_int32),
                ("device", C.c_int32), ("origin", C.c_float * 3), ("vox_dim", C.c_int32 * 3),
                ("color", C.c_int32), ("seq_batch", C.c_int32), ("carve", C.c_int32), ("reserved", C.c_int32 * 5)]


class Frame(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("depth_format", C.c_int32), ("depth_scale", C.c_float),
                ("pitch", C.c_int64), ("rgb", C.c_void_p), ("c2w", C.c_float * 16)]


class Stats(C.Structure):
    _fields_ = [("frames", C.c_int64), ("blocks_allocated", C.c_int64), ("pool_capacity", C.c_int64),
                ("slots_used", C.c_int64), ("touched_last", C.c_int64), ("touched_total", C.c_int64),
                ("new_total", C.c_int64), ("storage", C.c_int32), ("flags", C.c_int32),
                ("integrate_ms", C.c_double), ("alloc_ms", C.c_double), ("profiled_frames", C.c_int64),
                ("track_ms", C.c_double), ("profiled_launches", C.c_int64), ("blocks_integrated", C.c_int64),
                ("host_blocks", C.c_int64), ("streamed_out_total", C.c_int64), ("streamed_in_total", C.c_int64),
                ("stream_merged_total", C.c_int64), ("carved_total", C.c_int64), ("carve_ms", C.c_double),
                ("fused_ms", C.c_double), ("fused_launches", C.c_int64)]


class Mesh(C.Structure):
    _fields_ = [("n_tris", C.c_int64), ("xyz", C.POINTER(C.c_float)), ("cubes", C.POINTER(C.c_int32)),
                ("rgb", C.POINTER(C.c_uint8))]


class IndexedMesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("n_faces", C.c_int64), ("vertices", C.POINTER(C.c_float)),
                ("faces", C.POINTER(C.c_int32)), ("rgb", C.POINTER(C.c_uint8))]


class TrackConfig(C.Structure):
    _fields_ = [("d_prior_c2w", C.c_void_p), ("d_out_c2w", C.c_void_p), ("d_stats", C.c_void_p),
                ("gn_iters", C.c_int32), ("point_stride", C.c_int32), ("max_point_depth", C.c_float),
                ("prior_mode", C.c_int32), ("prior_sigma_t", C.c_float), ("prior_sigma_rot", C.c_float),
                ("depth_points", C.c_float), ("reserved", C.c_int32 * 1)]


class Settings(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("depth_factor", C.c_float), ("max_depth", C.c_float),
                ("origin", C.c_float * 3), ("vox_size", C.c_float), ("trunc", C.c_float), ("vox_dim", C.c_int32 * 3),
                ("found", C.c_uint32)]


class GnStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("n_valid", C.c_int32 * 64), ("rms", C.c_double * 64),
                ("step_norm", C.c_double * 64), ("pose", (C.c_double * 12) * 64)]


class SceneInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("n_frames", C.c_int32), ("fx", C.c_float),
                ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float), ("extent", C.c_float * 3),
                ("n_spheres", C.c_int32), ("n_boxes", C.c_int32), ("dropout", C.c_float)]


_lib = None

# (name, restype, argtypes) of every exported function
_PROTOS = [
    ("vh_config_default", None, [C.POINTER(Config)]),
    ("vh_create", C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    ("vh_destroy", C.c_int, [C.c_void_p]),
    ("vh_reset", C.c_int, [C.c_void_p]),
    ("vh_last_error", C.c_char_p, []),
    ("vh_abi_version", C.c_int, []),
    ("vh_integrate", C.c_int, [C.c_void_p, C.POINTER(Frame)]),
    ("vh_integrate_async", C.c_int, [C.c_void_p, C.POINTER(Frame)]),
    ("vh_bind_sequence", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_int32]),
    ("vh_bind_tracking", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_int32,
                                   C.POINTER(TrackConfig)]),
    ("vh_bind_sequence_rgb", C.c_int, [C.c_void_p, C.c_void_p]),
    ("vh_integrate_next", C.c_int, [C.c_void_p, C.c_int32]),
    ("vh_sync", C.c_int, [C.c_void_p]),
    ("vh_set_profiling", C.c_int, [C.c_void_p, C.c_int32]),
    ("vh_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("vh_stream", C.c_void_p, [C.c_void_p]),
    ("vh_extract_mesh", C.c_int, [C.c_void_p, C.POINTER(Mesh)]),
    ("vh_mesh_free", C.c_int, [C.POINTER(Mesh)]),
    ("vh_extract_mesh_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    ("vh_extract_mesh_device_rgb", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                             C.POINTER(C.c_int64)]),
    ("vh_extract_mesh_indexed_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                                 C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("vh_extract_mesh_indexed", C.c_int, [C.c_void_p, C.POINTER(IndexedMesh)]),
    ("vh_mc_tables", C.c_int, [C.c_void_p, C.c_void_p]),
    ("vh_indexed_mesh_free", C.c_int, [C.POINTER(IndexedMesh)]),
    ("vh_save_ply", C.c_int, [C.c_void_p, C.c_char_p]),
    ("vh_stream_out", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_int64)]),
    ("vh_stream_in", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_int64)]),
    ("vh_save_volume", C.c_int, [C.c_void_p, C.c_char_p]),
    ("vh_load_volume", C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("vh_get_config", C.c_int, [C.c_void_p, C.POINTER(Config)]),
    ("vh_export_blocks", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                   C.POINTER(C.c_int64)]),
    ("vh_export_block_colors", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    ("vh_query_sdf", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("vh_query_sdf_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("vh_query_sdf_backward_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("vh_gn_refine", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.POINTER(GnStats)]),
    ("vh_gn_refine_device_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                                              C.POINTER(GnStats)]),
    ("vh_pool_view", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    ("vh_settings_load", C.c_int, [C.c_char_p, C.POINTER(Settings)]),
    ("vh_settings_to_config", C.c_int, [C.POINTER(Settings), C.POINTER(Config)]),
    ("vh_png16_info", C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("vh_png16_load", C.c_int, [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32]),
    ("vh_png16_save", C.c_int, [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32]),
    ("vh_pose_load", C.c_int, [C.c_char_p, C.c_void_p]),
    ("vh_frame_load", C.c_int, [C.c_char_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    ("vh_scene_create", C.c_int, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("vh_scene_destroy", C.c_int, [C.c_void_p]),
    ("vh_scene_info_get", C.c_int, [C.c_void_p, C.POINTER(SceneInfo)]),
    ("vh_scene_pose", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("vh_scene_render_u16", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("vh_scene_render_clean", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("vh_scene_render_rgb", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("vh_scene_render_range", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
]


def _hip_runtimes_mapped():
    try:
        with open("/proc/self/maps") as f:
            paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    except OSError:
        return set()
    return {os.path.realpath(p) for p in paths}


def load():
    """Loads libvh.so (after torch, so that one HIP runtime serves both)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (binds the process to torch's HIP runtime first)
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise VhError(VH_ERR_DEVICE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in _PROTOS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    runtimes = _hip_runtimes_mapped()
    if len(runtimes) > 1:
        raise VhError(VH_ERR_DEVICE, f"two HIP runtimes are mapped ({sorted(runtimes)}); import torch first")
    _lib = lib
    return lib


def _check(rc):
    if rc != VH_OK:
        raise VhError(rc, _lib.vh_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _tptr(t):
    """Device pointer of a torch tensor (or None)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def make_config(**kw) -> Config:
    lib = load()
    c = Config()
    lib.vh_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def config_from_workload(w: dict, **over) -> Config:
    keys = ("width", "height", "fx", "fy", "cx", "cy", "max_depth", "min_depth", "vox_size", "trunc",
            "log2_buckets", "pool_blocks", "storage", "weight_cap", "color", "seq_batch", "carve")
    kw = {k: w[k] for k in keys if k in w}
    kw.update(over)
    return make_config(**kw)


class Volume:
    """One TSDF volume on one GPU (the engine behind ark::GpuTsdfGenerator)."""

    def __init__(self, cfg: Config, _handle=None):
        lib = load()
        self.cfg = cfg
        h = _handle
        if h is None:
            h = C.c_void_p()
            _check(lib.vh_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._pending = []  # host frames queued by integrate_async, referenced until sync()

    @classmethod
    def load(cls, path: str, device: int = -1) -> "Volume":
        """A volume from a checkpoint written by save() (vh_load_volume)."""
        lib = load()
        h = C.c_void_p()
        _check(lib.vh_load_volume(path.encode(), device, C.byref(h)))
        cfg = Config()
        _check(lib.vh_get_config(h, C.byref(cfg)))
        return cls(cfg, _handle=h)

    def save(self, path: str):
        """Checkpoint: configuration, frame counters, every block (vh_save_volume)."""
        _check(_lib.vh_save_volume(self._h, path.encode()))

    # -- out-of-core streaming
    def stream_out(self, center, radius: float) -> int:
        """Moves resident blocks farther than radius (metres) from center to the host store."""
        c = (C.c_float * 3)(*[float(x) for x in center])
        n = C.c_int64(0)
        _check(_lib.vh_stream_out(self._h, c, float(radius), C.byref(n)))
        return n.value

    def stream_in(self, center=(0.0, 0.0, 0.0), radius: float = float("inf")) -> int:
        """Moves host-store blocks within radius of center back into the pool."""
        c = (C.c_float * 3)(*[float(x) for x in center])
        n = C.c_int64(0)
        _check(_lib.vh_stream_in(self._h, c, float(radius), C.byref(n)))
        return n.value

    # -- lifetime
    def close(self):
        if self._h:
            _check(_lib.vh_destroy(self._h))
            self._h = None
            self._pending = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        _check(_lib.vh_reset(self._h))

    # -- integration
    @staticmethod
    def _rgb_arg(rgb, shape):
        """Host RGB image (H, W, 3) uint8 as a pointer and the array that owns it."""
        if rgb is None:
            return None, None
        if hasattr(rgb, "data_ptr"):
            import torch

            if rgb.dtype != torch.uint8 or rgb.is_cuda or not rgb.is_contiguous() or rgb.numel() != shape[0] * shape[1] * 3:
                raise TypeError("rgb tensor must be a contiguous CPU uint8 tensor of H*W*3 values")
            return C.c_void_p(rgb.data_ptr()), rgb
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        if rgb.size != shape[0] * shape[1] * 3:
            raise ValueError("rgb must hold H*W*3 bytes")
        return rgb.ctypes.data_as(C.c_void_p), rgb

    def integrate(self, depth: np.ndarray, c2w, scale: float = 0.001, rgb=None):
        """rgb: (H, W, 3) uint8, required by colour volumes."""
        depth = np.ascontiguousarray(depth)
        f = Frame()
        f.depth = depth.ctypes.data_as(C.c_void_p)
        if depth.dtype == np.uint16:
            f.depth_format = DEPTH_U16
            f.depth_scale = scale
        elif depth.dtype == np.float32:
            f.depth_format = DEPTH_F32
            f.depth_scale = 1.0
        else:
            raise TypeError("depth must be uint16 or float32")
        f.pitch = depth.strides[0]
        f.rgb, rgb_keep = self._rgb_arg(rgb, depth.shape)
        T = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32).reshape(16))
        for k in range(16):
            f.c2w[k] = float(T[k])
        _check(_lib.vh_integrate(self._h, C.byref(f)))

    def integrate_async(self, depth, c2w, scale: float = 0.001, rgb=None):
        """Queues one host frame (vh_integrate_async). depth: numpy uint16/float32 array or a
        CPU torch tensor (int16 holding uint16 bits, or float32), ideally page-locked
        (pin_memory=True); rgb likewise (uint8, H*W*3, colour volumes). Both are kept
        referenced until sync()."""
        f = Frame()
        if hasattr(depth, "data_ptr"):
            import torch

            if depth.dtype not in (torch.int16, torch.float32) or depth.is_cuda or not depth.is_contiguous():
                raise TypeError("depth tensor must be a contiguous CPU int16 (uint16 bits) or float32 tensor")
            f.depth = depth.data_ptr()
            u16 = depth.dtype == torch.int16
            f.pitch = depth.shape[-1] * (2 if u16 else 4)
        else:
            depth = np.ascontiguousarray(depth)
            if depth.dtype not in (np.uint16, np.float32):
                raise TypeError("depth must be uint16 or float32")
            f.depth = depth.ctypes.data_as(C.c_void_p)
            u16 = depth.dtype == np.uint16
            f.pitch = depth.strides[0]
        f.depth_format = DEPTH_U16 if u16 else DEPTH_F32
        f.depth_scale = scale if u16 else 1.0
        f.rgb, rgb_keep = self._rgb_arg(rgb, (self.cfg.height, self.cfg.width))
        T = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32).reshape(16))
        for k in range(16):
            f.c2w[k] = float(T[k])
        _check(_lib.vh_integrate_async(self._h, C.byref(f)))
        self._pending.append((depth, rgb_keep))

    def bind_sequence(self, d_depth, fmt: int, scale: float, d_c2w, n: int):
        """d_depth / d_c2w: torch device tensors (or integer device addresses)."""
        dp = d_depth if isinstance(d_depth, int) else d_depth.data_ptr()
        pp = d_c2w if isinstance(d_c2w, int) else d_c2w.data_ptr()
        _check(_lib.vh_bind_sequence(self._h, C.c_void_p(dp), fmt, scale, C.c_void_p(pp), n))

    def bind_tracking(self, d_depth, fmt: int, scale: float, n: int, d_prior, d_out, d_stats=None,
                      gn_iters: int = 5, point_stride: int = 4, max_point_depth: float = 0.0,
                      prior_mode: int = PRIOR_ABSOLUTE, prior_sigma_t: float = 0.0, prior_sigma_rot: float = 0.0,
                      depth_points: float = 0.0):
        """Tracked sequence: per frame, gn_iters solver steps from its starting pose (d_prior[f],
        or with PRIOR_INCREMENT the refined pose of frame f-1 times d_prior[f]) over the
        frame's back-projected depth (points up to max_point_depth, 0 = the volume's
        max_depth), then integration at the refined pose (written to d_out[f]).
        d_stats: None or a float64 device tensor of n * (gn_iters + 1) * GN_STATS values.
        prior_sigma_t (metres) / prior_sigma_rot (radians) > 0 add a motion prior centred on
        the starting pose; depth_points > 0 scales the depth term to that many points."""
        tc = TrackConfig()
        tc.d_prior_c2w = d_prior.data_ptr()
        tc.d_out_c2w = d_out.data_ptr()
        tc.d_stats = None if d_stats is None else d_stats.data_ptr()
        tc.gn_iters, tc.point_stride = gn_iters, point_stride
        tc.max_point_depth, tc.prior_mode = max_point_depth, prior_mode
        tc.prior_sigma_t, tc.prior_sigma_rot, tc.depth_points = prior_sigma_t, prior_sigma_rot, depth_points
        _check(_lib.vh_bind_tracking(self._h, _tptr(d_depth), fmt, scale, n, C.byref(tc)))

    def bind_sequence_rgb(self, d_rgb):
        """RGB frames (uint8 device tensor, n * H * W * 3) of the sequence bound last."""
        _check(_lib.vh_bind_sequence_rgb(self._h, _tptr(d_rgb)))

    def integrate_next(self, n: int = 1):
        _check(_lib.vh_integrate_next(self._h, n))

    def sync(self):
        try:
            _check(_lib.vh_sync(self._h))
        finally:
            self._pending = []

    def set_profiling(self, on):
        """False/0: off; True/1: eager launches with events around each kernel; 2: the captured
        graphs with events around each fused batch launch (stats fused_ms, fused_launches)."""
        _check(_lib.vh_set_profiling(self._h, 2 if on == 2 else (1 if on else 0)))

    def stats(self) -> dict:
        s = Stats()
        _check(_lib.vh_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def stream(self) -> int:
        return _lib.vh_stream(self._h) or 0

    # -- blocks / mesh
    def export_blocks(self):
        """(keys (N,3) int32, tsdf (N,512) float32, weight (N,512) float32), pool order."""
        n = C.c_int64(0)
        rc = _lib.vh_export_blocks(self._h, None, None, None, 0, C.byref(n))
        if rc not in (VH_OK, VH_ERR_TRUNCATED):
            _check(rc)
        nb = n.value
        keys = np.zeros((nb, 3), np.int32)
        tsdf = np.zeros((nb, 512), np.float32)
        w = np.zeros((nb, 512), np.float32)
        _check(_lib.vh_export_blocks(self._h, _ptr(keys), _ptr(tsdf), _ptr(w), nb, C.byref(n)))
        return keys, tsdf, w

    def export_block_colors(self):
        """(N, 512, 3) uint8 voxel colours in export_blocks' order (colour volumes)."""
        n = C.c_int64(0)
        rc = _lib.vh_export_block_colors(self._h, None, 0, C.byref(n))
        if rc not in (VH_OK, VH_ERR_TRUNCATED):
            _check(rc)
        nb = n.value
        rgb = np.zeros((nb, 512, 3), np.uint8)
        _check(_lib.vh_export_block_colors(self._h, _ptr(rgb), nb, C.byref(n)))
        return rgb

    def extract_mesh(self, with_rgb: bool = False):
        """(xyz (T,9) float32 metres, cubes (T,4) int32) and, with_rgb, rgb (T,9) uint8 (None
        for a volume without colour)."""
        m = Mesh()
        _check(_lib.vh_extract_mesh(self._h, C.byref(m)))
        try:
            n = m.n_tris
            xyz = np.ctypeslib.as_array(m.xyz, shape=(max(n, 1) * 9,))[: n * 9].reshape(n, 9).copy()
            cubes = np.ctypeslib.as_array(m.cubes, shape=(max(n, 1) * 4,))[: n * 4].reshape(n, 4).copy()
            rgb = None
            if m.rgb:
                rgb = np.ctypeslib.as_array(m.rgb, shape=(max(n, 1) * 9,))[: n * 9].reshape(n, 9).copy()
        finally:
            _lib.vh_mesh_free(C.byref(m))
        return (xyz, cubes, rgb) if with_rgb else (xyz, cubes)

    def extract_mesh_device(self, d_xyz, d_cubes, capacity: int, d_rgb=None) -> int:
        n = C.c_int64(0)
        rc = _lib.vh_extract_mesh_device_rgb(self._h, _tptr(d_xyz), _tptr(d_cubes), _tptr(d_rgb), capacity,
                                             C.byref(n))
        if rc not in (VH_OK, VH_ERR_TRUNCATED):
            _check(rc)
        return n.value

    def extract_mesh_indexed(self, with_rgb: bool = False):
        """(vertices (V,3) float32 metres, faces (F,3) int32) and, with_rgb, rgb (V,3) uint8 (None
        without colour): one vertex per crossed lattice edge, shared by its faces; the faces are
        every table triangle, zero-area ones included, so a closed surface gives a closed mesh."""
        m = IndexedMesh()
        _check(_lib.vh_extract_mesh_indexed(self._h, C.byref(m)))
        try:
            nv, nf = m.n_vertices, m.n_faces
            v = np.ctypeslib.as_array(m.vertices, shape=(max(nv, 1) * 3,))[: nv * 3].reshape(nv, 3).copy()
            f = np.ctypeslib.as_array(m.faces, shape=(max(nf, 1) * 3,))[: nf * 3].reshape(nf, 3).copy()
            rgb = None
            if m.rgb:
                rgb = np.ctypeslib.as_array(m.rgb, shape=(max(nv, 1) * 3,))[: nv * 3].reshape(nv, 3).copy()
        finally:
            _lib.vh_indexed_mesh_free(C.byref(m))
        return (v, f, rgb) if with_rgb else (v, f)

    def extract_mesh_indexed_device(self, d_vertices, vertex_capacity: int, d_faces, face_capacity: int,
                                    d_vrgb=None):
        """Device variant; returns (n_vertices, n_faces, status)."""
        nv, nf = C.c_int64(0), C.c_int64(0)
        rc = _lib.vh_extract_mesh_indexed_device(self._h, _tptr(d_vertices), vertex_capacity, _tptr(d_faces),
                                                 face_capacity, _tptr(d_vrgb), C.byref(nv), C.byref(nf))
        if rc not in (VH_OK, VH_ERR_TRUNCATED):
            _check(rc)
        return nv.value, nf.value, rc

    def save_ply(self, path: str):
        _check(_lib.vh_save_ply(self._h, path.encode()))

    # -- queries
    def query(self, pts: np.ndarray):
        pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        n = pts.shape[0]
        sdf = np.zeros(n, np.float32)
        valid = np.zeros(n, np.uint8)
        _check(_lib.vh_query_sdf(self._h, _ptr(pts), n, _ptr(sdf), _ptr(valid)))
        return sdf, valid.astype(bool)

    def query_device(self, d_pts, n, d_sdf, d_valid):
        _check(_lib.vh_query_sdf_device(self._h, _tptr(d_pts), n, _tptr(d_sdf), _tptr(d_valid)))

    def backward_device(self, d_pts, n, c2w, d_grad_r=None, d_r=None, d_valid=None, d_J=None, d_pose_grad=None,
                        d_voxel_grad=None):
        T = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32).reshape(16))
        _check(_lib.vh_query_sdf_backward_device(self._h, _tptr(d_pts), n, _ptr(T), _tptr(d_grad_r), _tptr(d_r),
                                                 _tptr(d_valid), _tptr(d_J), _tptr(d_pose_grad),
                                                 _tptr(d_voxel_grad)))

    def gn_refine(self, pts_cam: np.ndarray, c2w, iters: int):
        pts = np.ascontiguousarray(pts_cam, dtype=np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(np.asarray(c2w, dtype=np.float64).reshape(16)).copy()
        st = GnStats()
        _check(_lib.vh_gn_refine(self._h, _ptr(pts), pts.shape[0], _ptr(T), iters, C.byref(st)))
        return T.reshape(4, 4), self._gn_stats(st, iters)

    def gn_refine_device_noalloc(self, d_pts, n: int, c2w, iters: int):
        """vh_gn_refine_device_noalloc: the refinement over device points (float32 tensor of
        n x 3), with the solver scratch in the library's static device array."""
        T = np.ascontiguousarray(np.asarray(c2w, dtype=np.float64).reshape(16)).copy()
        st = GnStats()
        _check(_lib.vh_gn_refine_device_noalloc(self._h, _tptr(d_pts), n, _ptr(T), iters, C.byref(st)))
        return T.reshape(4, 4), self._gn_stats(st, iters)

    @staticmethod
    def _gn_stats(st, iters):
        k = iters + 1
        return {"n_valid": list(st.n_valid[:k]), "rms": list(st.rms[:k]), "step_norm": list(st.step_norm[:k]),
                "pose": [np.vstack([np.array(st.pose[i][:], np.float64).reshape(3, 4), [0, 0, 0, 1]])
                         for i in range(k)]}

    def pool_view(self):
        vox = C.c_void_p()
        keys = C.c_void_p()
        n = C.c_int64()
        _check(_lib.vh_pool_view(self._h, C.byref(vox), C.byref(keys), C.byref(n)))
        return vox.value, keys.value, n.value


def load_settings(path: str) -> dict:
    """vh_settings_load: the settings file's camera and voxel keys (missing keys absent)."""
    st = Settings()
    _check(load().vh_settings_load(os.fsencode(path), C.byref(st)))
    out = {"found": st.found}
    bits = [("fx", 1), ("fy", 2), ("cx", 4), ("cy", 8), ("width", 16), ("height", 32), ("depth_factor", 64),
            ("max_depth", 128), ("origin", 256), ("vox_size", 512), ("trunc", 1024), ("vox_dim", 2048)]
    for name, bit in bits:
        if st.found & bit:
            v = getattr(st, name)
            out[name] = tuple(v) if name in ("origin", "vox_dim") else v
    return out


def config_from_settings(path: str, **over) -> Config:
    """vh_config with the camera and voxel fields of a settings file (vh_settings_to_config)."""
    lib = load()
    st = Settings()
    _check(lib.vh_settings_load(os.fsencode(path), C.byref(st)))
    cfg = make_config()
    _check(lib.vh_settings_to_config(C.byref(st), C.byref(cfg)))
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg


def png16_load(path: str) -> np.ndarray:
    lib = load()
    w, h = C.c_int32(), C.c_int32()
    _check(lib.vh_png16_info(os.fsencode(path), C.byref(w), C.byref(h)))
    out = np.zeros((h.value, w.value), np.uint16)
    _check(lib.vh_png16_load(os.fsencode(path), _ptr(out), w.value, h.value))
    return out


def png16_save(path: str, img: np.ndarray):
    img = np.ascontiguousarray(img, dtype=np.uint16)
    _check(load().vh_png16_save(os.fsencode(path), _ptr(img), img.shape[1], img.shape[0]))


def pose_load(path: str) -> np.ndarray:
    T = np.zeros(16, np.float32)
    _check(load().vh_pose_load(os.fsencode(path), _ptr(T)))
    return T.reshape(4, 4)


def frame_load(directory: str, frame_id: int, width: int, height: int):
    """(depth uint16 (height, width), c2w (4, 4) float32) of <dir>/depth/<id>.png, <dir>/tcw/<id>.txt."""
    d = np.zeros((height, width), np.uint16)
    T = np.zeros(16, np.float32)
    _check(load().vh_frame_load(os.fsencode(directory), frame_id, _ptr(d), width, height, _ptr(T)))
    return d, T.reshape(4, 4)


class Scene:
    """Seeded synthetic depth sequence (include/vh_synth.h)."""

    def __init__(self, preset: int, seed: int):
        lib = load()
        h = C.c_void_p()
        rc = lib.vh_scene_create(preset, seed, C.byref(h))
        if rc:
            raise VhError(rc, f"vh_scene_create({preset}) failed")
        self._h = h
        self._lib = lib  # held for __del__: module globals may already be cleared at exit
        info = SceneInfo()
        lib.vh_scene_info_get(h, C.byref(info))
        self.width, self.height, self.n_frames = info.width, info.height, info.n_frames
        self.intr = (info.fx, info.fy, info.cx, info.cy)
        self.extent = tuple(info.extent)
        self.n_spheres, self.n_boxes = info.n_spheres, info.n_boxes

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.vh_scene_destroy(self._h)
            self._h = None

    def pose(self, f: int) -> np.ndarray:
        T = np.zeros(16, np.float32)
        _check(_lib.vh_scene_pose(self._h, f, _ptr(T)))
        return T.reshape(4, 4)

    def depth(self, f: int, c2w=None) -> np.ndarray:
        d = np.zeros((self.height, self.width), np.uint16)
        Tp = None if c2w is None else np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        _check(_lib.vh_scene_render_u16(self._h, f, _ptr(Tp), _ptr(d)))
        return d

    def clean_depth(self, f: int, c2w=None) -> np.ndarray:
        d = np.zeros((self.height, self.width), np.float32)
        Tp = None if c2w is None else np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        _check(_lib.vh_scene_render_clean(self._h, f, _ptr(Tp), _ptr(d)))
        return d

    def rgb(self, f: int, c2w=None) -> np.ndarray:
        """Colour image (H, W, 3) uint8 of frame f (or of pose c2w)."""
        out = np.zeros((self.height, self.width, 3), np.uint8)
        Tp = None if c2w is None else np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        _check(_lib.vh_scene_render_rgb(self._h, f, _ptr(Tp), _ptr(out)))
        return out

    def frames(self, f0: int, n: int):
        """(depth (n,H,W) uint16, c2w (n,4,4) float32)."""
        d = np.zeros((n, self.height, self.width), np.uint16)
        T = np.zeros((n, 16), np.float32)
        _check(_lib.vh_scene_render_range(self._h, f0, n, _ptr(d), _ptr(T)))
        return d, T.reshape(n, 4, 4)
