"""ctypes loader of the CPU oracle (oracle/libvh_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
CPU-baseline leg of bench.py, always as the oracle that libvh's results are compared with,
never as the thing measured or shipped. See vh_oracle.cpp for the reference lines each function restates.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvh_oracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.vho_create.restype = P
    lib.vho_create.argtypes = [C.c_int, C.c_int] + [C.c_float] * 8 + [C.c_int, C.c_int]
    lib.vho_destroy.argtypes = [P]
    lib.vho_set_threads.argtypes = [C.c_int]
    lib.vho_max_threads.restype = C.c_int
    lib.vho_integrate.restype = C.c_int64
    lib.vho_integrate.argtypes = [P, P, C.c_float, P]
    lib.vho_num_blocks.restype = C.c_int64
    lib.vho_num_blocks.argtypes = [P]
    lib.vho_set_carve.argtypes = [P, C.c_int]
    lib.vho_set_shadow.argtypes = [P, C.c_int]
    lib.vho_scatter_grad.restype = C.c_int64
    lib.vho_scatter_grad.argtypes = [P, P, P, C.c_int64, P, C.c_int]
    lib.vho_touched.restype = C.c_int64
    lib.vho_touched.argtypes = [P, P, C.c_int64]
    lib.vho_export.restype = C.c_int64
    lib.vho_export.argtypes = [P, P, P, P, P, P, C.c_int64]
    lib.vho_mc.restype = C.c_int64
    lib.vho_mc.argtypes = [P, P, P, C.c_int64]
    lib.vho_mc_cube.restype = C.c_int
    lib.vho_mc_cube.argtypes = [P, P]
    lib.vho_tri_case.restype = C.c_char_p
    lib.vho_tri_case.argtypes = [C.c_int]
    lib.vho_query.argtypes = [P, P, C.c_int64, P, P, P, P]
    lib.vho_query64.argtypes = [P, P, C.c_int64, P, P, P]
    lib.vho_query_twin.argtypes = [P, P, C.c_int64, P, P, P]
    lib.vho_band.restype = C.c_int
    lib.vho_band.argtypes = [P, P, C.c_int, C.c_int, C.c_float, P, C.c_int]
    lib.vho_f32_to_f16.restype = C.c_uint16
    lib.vho_f32_to_f16.argtypes = [C.c_float]
    lib.vho_f16_to_f32.restype = C.c_float
    lib.vho_f16_to_f32.argtypes = [C.c_uint16]
    lib.vho_vertex_interp.argtypes = [P, P, C.c_float, C.c_float, P]
    lib.vho_set_color.restype = C.c_int
    lib.vho_set_color.argtypes = [P, C.c_int]
    lib.vho_integrate_rgb.restype = C.c_int64
    lib.vho_integrate_rgb.argtypes = [P, P, P, C.c_float, P]
    lib.vho_export_colors.restype = C.c_int64
    lib.vho_export_colors.argtypes = [P, P, C.c_int64]
    lib.vho_mc_rgb.restype = C.c_int64
    lib.vho_mc_rgb.argtypes = [P, P, P, P, C.c_int64]
    lib.vho_vertex_interp_rgb.argtypes = [P, P, C.c_float, C.c_float, P, P, P, P]
    lib.vho_color_update.argtypes = [P, C.c_float, P]
    lib.vho_mc64.restype = C.c_int64
    lib.vho_mc64.argtypes = [P, P, P, C.c_int64]
    lib.vho_point_triangle_dist.restype = C.c_double
    lib.vho_point_triangle_dist.argtypes = [P, P, P, P]
    lib.vho_mesh_distance.argtypes = [P, P, C.c_int64, P, P, C.c_int64, P, P]
    lib.vho_import_f32.restype = C.c_int
    lib.vho_import_f32.argtypes = [P, P, P, P, C.c_int64]
    lib.vho_mesh_indexed.restype = C.c_int64
    lib.vho_mesh_indexed.argtypes = [P, P, P, C.c_int64, P, P, C.c_int64, P]
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def sort_blocks(keys, *arrays):
    """Sorts blocks by (x, y, z); returns the sorted keys and arrays."""
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    return (keys[order],) + tuple(a[order] for a in arrays)


class Oracle:
    """color: a colour volume (integrate then needs rgb); carve: every allocated voxel a frame
    observes is updated (default: the workload's "carve" key, else off); shadow=False: the FP64
    shadow is not kept (the FP32 twin alone, for timing)."""

    def __init__(self, w: dict, storage=None, weight_cap=None, threads=0, color=False, carve=None, shadow=True):
        lib = load()
        if threads:
            lib.vho_set_threads(threads)
        st = w["storage"] if storage is None else storage
        cap = w["weight_cap"] if weight_cap is None else weight_cap
        self.w = dict(w)
        self._h = lib.vho_create(w["width"], w["height"], w["fx"], w["fy"], w["cx"], w["cy"], w["max_depth"],
                                 w["min_depth"], w["vox_size"], w["trunc"], st, cap)
        self.color = bool(color)
        if self.color:
            lib.vho_set_color(self._h, 1)
        self.carve = int(w.get("carve", 0) if carve is None else carve)  # 2: no culling bound
        lib.vho_set_carve(self._h, self.carve)
        if not shadow:
            lib.vho_set_shadow(self._h, 0)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.vho_destroy(self._h)
            self._h = None

    def integrate(self, depth_u16: np.ndarray, c2w, scale=0.001, rgb=None) -> int:
        d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
        T = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        if self.color:
            c = np.ascontiguousarray(rgb, dtype=np.uint8)
            assert c.size == d.size * 3
            return int(_lib.vho_integrate_rgb(self._h, _p(d), _p(c), scale, _p(T)))
        return int(_lib.vho_integrate(self._h, _p(d), scale, _p(T)))

    def num_blocks(self) -> int:
        return int(_lib.vho_num_blocks(self._h))

    def touched(self) -> np.ndarray:
        n = _lib.vho_touched(self._h, None, 0)
        out = np.zeros((n, 3), np.int32)
        _lib.vho_touched(self._h, _p(out), n)
        return out

    def export(self):
        """Sorted (keys, tsdf, w, tsdf64, w64)."""
        n = self.num_blocks()
        keys = np.zeros((n, 3), np.int32)
        s = np.zeros((n, 512), np.float32)
        w = np.zeros((n, 512), np.float32)
        s64 = np.zeros((n, 512), np.float64)
        w64 = np.zeros((n, 512), np.float64)
        _lib.vho_export(self._h, _p(keys), _p(s), _p(w), _p(s64), _p(w64), n)
        return sort_blocks(keys, s, w, s64, w64)

    def scatter_grad(self, pts64, up, twin=False):
        """Voxel gradient of sum_i up_i r_i (CPU scatter, FP64 sums): (keys, grad (N, 512))
        sorted like export(), and the number of points with a valid sample. twin: FP32
        weights and products as the kernels form them (FP64 weights otherwise)."""
        n = self.num_blocks()
        keys = np.zeros((n, 3), np.int32)
        _lib.vho_export(self._h, _p(keys), None, None, None, None, n)
        pts = np.ascontiguousarray(pts64, np.float64)
        u = np.ascontiguousarray(up, np.float32)
        g = np.zeros((n, 512), np.float64)
        used = _lib.vho_scatter_grad(self._h, _p(pts), _p(u), len(pts), _p(g), int(bool(twin)))
        k, g = sort_blocks(keys, g)
        return k, g, int(used)

    def export_rgb(self):
        """Sorted (keys, rgb (N, 512, 3) uint8) of a colour volume."""
        n = self.num_blocks()
        keys = np.zeros((n, 3), np.int32)
        _lib.vho_export(self._h, _p(keys), None, None, None, None, n)
        rgb = np.zeros((n, 512, 3), np.uint8)
        _lib.vho_export_colors(self._h, _p(rgb), n)
        return sort_blocks(keys, rgb)

    def mc(self, with_rgb=False):
        n = _lib.vho_mc(self._h, None, None, 0)
        xyz = np.zeros((n, 9), np.float32)
        cubes = np.zeros((n, 4), np.int32)
        rgb = np.zeros((n, 9), np.uint8) if with_rgb else None
        _lib.vho_mc_rgb(self._h, _p(xyz), _p(cubes), _p(rgb), n)
        return (xyz, cubes, rgb) if with_rgb else (xyz, cubes)

    def mc64(self):
        """The FP64 reference mesh (marching cubes over the FP64 shadow): (xyz, cubes)."""
        n = _lib.vho_mc64(self._h, None, None, 0)
        xyz = np.zeros((n, 9), np.float32)
        cubes = np.zeros((n, 4), np.int32)
        _lib.vho_mc64(self._h, _p(xyz), _p(cubes), n)
        return xyz, cubes

    def import_f32(self, keys, tsdf, weight):
        """Blocks with chosen values (keys (N,3) block coordinates, tsdf / weight (N,512)) into
        an FP32 twin."""
        keys = np.ascontiguousarray(keys, np.int32)
        tsdf = np.ascontiguousarray(tsdf, np.float32)
        weight = np.ascontiguousarray(weight, np.float32)
        assert _lib.vho_import_f32(self._h, _p(keys), _p(tsdf), _p(weight), len(keys)) == 0

    def mesh_indexed(self, with_rgb=False, with_cubes=False):
        """(vertices (V,3) float32 metres, faces (F,3) int32[, rgb (V,3) uint8][, cubes (F,4)
        int32]): every table triangle of the twin's observed cubes (zero-area ones included),
        one vertex per lattice edge they use, interpolated lower to upper end point."""
        nf = C.c_int64(0)
        nv = _lib.vho_mesh_indexed(self._h, None, None, 0, None, None, 0, C.cast(C.pointer(nf), C.c_void_p))
        nf = nf.value
        v = np.zeros((nv, 3), np.float32)
        f = np.zeros((nf, 3), np.int32)
        rgb = np.zeros((nv, 3), np.uint8) if with_rgb else None
        cubes = np.zeros((nf, 4), np.int32) if with_cubes else None
        _lib.vho_mesh_indexed(self._h, _p(v), _p(rgb), nv, _p(f), _p(cubes), nf, None)
        out = (v, f) + ((rgb,) if with_rgb else ()) + ((cubes,) if with_cubes else ())
        return out

    def query(self, pts):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        n = pts.shape[0]
        s32 = np.zeros(n, np.float32)
        s64 = np.zeros(n, np.float64)
        g64 = np.zeros((n, 3), np.float64)
        valid = np.zeros(n, np.uint8)
        _lib.vho_query(self._h, _p(pts), n, _p(s32), _p(s64), _p(g64), _p(valid))
        return s32, s64, g64, valid.astype(bool)

    def query64(self, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        s = np.zeros(n, np.float64)
        g = np.zeros((n, 3), np.float64)
        valid = np.zeros(n, np.uint8)
        _lib.vho_query64(self._h, _p(pts), n, _p(s), _p(g), _p(valid))
        return s, g, valid.astype(bool)

    def query_twin(self, pts64):
        """FP32 value and FP32 gradient as the GPU kernels compute them, at FP64 points."""
        pts = np.ascontiguousarray(pts64, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        s = np.zeros(n, np.float32)
        g = np.zeros((n, 3), np.float32)
        valid = np.zeros(n, np.uint8)
        _lib.vho_query_twin(self._h, _p(pts), n, _p(s), _p(g), _p(valid))
        return s, g, valid.astype(bool)

    def band(self, c2w, u, v, d):
        T = np.ascontiguousarray(np.asarray(c2w, np.float32).reshape(16))
        out = np.zeros((64, 3), np.int32)
        n = _lib.vho_band(self._h, _p(T), u, v, d, _p(out), 64)
        return out[: min(n, 64)]


def mc_cube(vals):
    load()
    v = np.ascontiguousarray(vals, np.float32)
    out = np.zeros((5, 9), np.float32)
    n = _lib.vho_mc_cube(_p(v), _p(out))
    return out[:n]


def tri_case(ci: int) -> str:
    return load().vho_tri_case(ci).decode()


def f32_to_f16_bits(x: float) -> int:
    return int(load().vho_f32_to_f16(x))


def f16_bits_to_f32(h: int) -> float:
    return float(load().vho_f16_to_f32(h))


def vertex_interp(a, b, va, vb):
    load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(3, np.float32)
    _lib.vho_vertex_interp(_p(a), _p(b), va, vb, _p(out))
    return out


def vertex_interp_rgb(a, b, va, vb, ca, cb):
    """(position, colour) of the MC vertex between corners a and b with colours ca, cb."""
    load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ca = np.ascontiguousarray(ca, np.uint8)
    cb = np.ascontiguousarray(cb, np.uint8)
    out = np.zeros(3, np.float32)
    cout = np.zeros(3, np.uint8)
    _lib.vho_vertex_interp_rgb(_p(a), _p(b), va, vb, _p(ca), _p(cb), _p(out), _p(cout))
    return out, cout


def point_triangle_dist(p, a, b, c):
    load()
    arr = [np.ascontiguousarray(x, np.float64) for x in (p, a, b, c)]
    return float(_lib.vho_point_triangle_dist(*[_p(x) for x in arr]))


def mesh_distance(a_xyz, a_cubes, b_xyz, b_cubes, quantiles=(0.999, 0.99999)):
    """One-sided distances from mesh A to mesh B (metres): dict(max, mean, far, samples) and
    the given quantiles of the per-sample distances (samples with a B triangle nearby)."""
    load()
    ax = np.ascontiguousarray(a_xyz, np.float32)
    ac = np.ascontiguousarray(a_cubes, np.int32)
    bx = np.ascontiguousarray(b_xyz, np.float32)
    bc = np.ascontiguousarray(b_cubes, np.int32)
    out = np.zeros(4, np.float64)
    dist = np.zeros(4 * len(ax), np.float32) if quantiles else None
    _lib.vho_mesh_distance(_p(ax), _p(ac), len(ax), _p(bx), _p(bc), len(bx), _p(out), _p(dist))
    res = {"max": float(out[0]), "mean": float(out[1]), "far": int(out[2]), "samples": int(out[3])}
    if quantiles:
        d = dist[dist >= 0]
        for q in quantiles:
            res[f"q{q}"] = float(np.quantile(d, q)) if len(d) else 0.0
    return res


def hausdorff(a_xyz, a_cubes, b_xyz, b_cubes, quantiles=(0.999, 0.99999)):
    """Symmetric distance between two marching-cubes meshes on the same lattice: the larger of
    the two one-sided maxima (samples: vertices and centroids), with both directions and the
    larger of their quantiles."""
    ab = mesh_distance(a_xyz, a_cubes, b_xyz, b_cubes, quantiles)
    ba = mesh_distance(b_xyz, b_cubes, a_xyz, a_cubes, quantiles)
    out = {"hausdorff": max(ab["max"], ba["max"]), "a_to_b": ab, "b_to_a": ba}
    for q in quantiles or ():
        out[f"q{q}"] = max(ab[f"q{q}"], ba[f"q{q}"])
    return out


def color_update(c, w_old, px):
    """One colour observation: stored colour c (3 bytes), stored weight w_old, pixel px."""
    load()
    c = np.array(c, np.uint8)
    px = np.ascontiguousarray(px, np.uint8)
    _lib.vho_color_update(_p(c), w_old, _p(px))
    return c
