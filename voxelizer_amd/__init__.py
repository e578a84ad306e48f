"""voxelizer_amd: MI355X-native voxelizer (hand-written HIP for gfx950).

Python binding of the C ABI in include/voxelizer_amd.h (libvoxelizer_amd.so,
built in-tree by voxelizer_amd/build.py).  There is no CPU fallback: if the
library or a HIP device is missing, the calls raise.

    import voxelizer_amd as va
    ctx = va.Context()
    ctx.set_mesh(vertices, indices)
    ctx.set_materials(materials, n_unique)
    grid = ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.01)
    nodes = ctx.voxelize_nodes(grid, va.NODE_LONG)        # numpy, copied back
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoxelizer_amd.so")

OK, ERR_INVALID, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_NO_DEVICE = range(6)
NODE_SHORT, NODE_LONG, NODE_PARTIAL, NODE_SHORT_FCC, NODE_LONG_FCC = range(5)
PLAN_MAX_DIMENSION, PLAN_CUBE_LENGTH = 0, 1
NODE_SIZE = {NODE_SHORT: 1, NODE_LONG: 2, NODE_PARTIAL: 8, NODE_SHORT_FCC: 2, NODE_LONG_FCC: 4}
DEFAULT_VOX_SPLIT = (1024, 512, 512)
DEFAULT_MAT_SPLIT = (512, 512, 512)
SCHEDULE_PLAIN, SCHEDULE_LEARNED = 0, 1
COST_FEATURES, COST_PARAMS = 16, 5313


class VoxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"vox status {status}: {msg}")
        self.status = status


class Grid(C.Structure):
    """vox_grid of include/voxelizer_amd.h."""
    _fields_ = [("min_vertex", C.c_double * 3), ("max_vertex", C.c_double * 3),
                ("voxel_length", C.c_double), ("res", C.c_uint32 * 3),
                ("alloc", C.c_uint32 * 3), ("x_splits", C.c_uint32), ("pad_", C.c_uint32)]

    def as_dict(self):
        return {"min_vertex": list(self.min_vertex), "max_vertex": list(self.max_vertex),
                "voxel_length": self.voxel_length, "res": list(self.res),
                "alloc": list(self.alloc), "x_splits": self.x_splits}

    def same(self, other) -> bool:
        return self.as_dict() == other.as_dict()


class BatchItem(C.Structure):
    """vox_batch_item of include/voxelizer_amd.h."""
    _fields_ = [("vertices_dev", C.c_void_p), ("nv", C.c_uint32), ("nt", C.c_uint32),
                ("indices_dev", C.c_void_p), ("materials_dev", C.c_void_p),
                ("n_unique", C.c_uint32), ("kind", C.c_int), ("grid", Grid),
                ("vox_split", C.c_uint32 * 3), ("mat_split", C.c_uint32 * 3),
                ("nodes_dev", C.c_void_p), ("status", C.c_int)]


class MeshC(C.Structure):
    """vox_mesh."""
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("nv", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("nt", C.c_uint32)]


class StageTimes(C.Structure):
    _fields_ = [("plan_ms", C.c_float), ("solid_bin_ms", C.c_float), ("solid_fill_ms", C.c_float),
                ("surface_ms", C.c_float), ("erode_ms", C.c_float), ("nodes_ms", C.c_float),
                ("total_ms", C.c_float), ("erosion_passes", C.c_uint32),
                ("solid_pairs", C.c_uint32), ("surface_pairs", C.c_uint32),
                ("surface_voxels", C.c_uint32), ("learned_schedule", C.c_uint32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


# Every symbol include/voxelizer_amd.h declares (checked by the CPU tests).
EXPORTED = [
    "vox_create", "vox_destroy", "vox_last_error", "vox_device_count", "vox_sync",
    "vox_build_info", "vox_alloc", "vox_free", "vox_memcpy_h2d", "vox_memcpy_d2h",
    "vox_set_mesh", "vox_set_mesh_device", "vox_set_materials", "vox_set_materials_device",
    "vox_plan_grid", "vox_plan_grid_host", "vox_solid_bits", "vox_voxelize_nodes",
    "vox_voxelize_nodes_noalloc", "vox_voxelize_partition", "vox_solid_bits_partition",
    "vox_voxelize_slice", "vox_node_dims",
    "vox_node_count", "vox_node_size", "vox_enable_timing", "vox_get_stage_times",
    "vox_set_schedule", "vox_set_heavy_threshold", "vox_set_screen_ratio", "vox_set_split_parts", "vox_cost_set_params",
    "vox_cost_get_params", "vox_cost_record",
    "vox_cost_samples", "vox_cost_train", "vox_mlp_forward", "vox_mlp_grad",
    "vox_voxelize_batch", "vox_set_brick_dims", "vox_get_brick_dims", "vox_autotune_bricks",
    "vox_graph_create", "vox_graph_launch", "vox_graph_fallbacks", "vox_graph_destroy",
    "vox_mesh_load", "vox_mesh_free", "vox_write_nodes_vtk", "vox_write_bits_vtk",
]

_lib = None


def lib():
    """Loads libvoxelizer_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m voxelizer_amd.build`")
    L = C.CDLL(LIB_PATH)
    P, U32, I32, D = C.c_void_p, C.c_uint32, C.c_int, C.c_double
    S = C.c_int
    L.vox_create.argtypes = [I32, C.POINTER(P)]
    L.vox_create.restype = S
    L.vox_destroy.argtypes = [P]
    L.vox_destroy.restype = None
    L.vox_last_error.argtypes = [P]
    L.vox_last_error.restype = C.c_char_p
    L.vox_device_count.restype = I32
    L.vox_sync.argtypes = [P]
    L.vox_build_info.restype = C.c_char_p
    L.vox_alloc.argtypes = [P, C.c_size_t, C.POINTER(P)]
    L.vox_free.argtypes = [P, P]
    L.vox_memcpy_h2d.argtypes = [P, P, P, C.c_size_t]
    L.vox_memcpy_d2h.argtypes = [P, P, P, C.c_size_t]
    L.vox_set_mesh.argtypes = [P, P, U32, P, U32]
    L.vox_set_mesh_device.argtypes = [P, P, U32, P, U32]
    L.vox_set_materials.argtypes = [P, P, U32]
    L.vox_set_materials_device.argtypes = [P, P, U32]
    L.vox_plan_grid.argtypes = [P, I32, D, U32, C.POINTER(Grid)]
    L.vox_plan_grid_host.argtypes = [P, U32, I32, D, U32, C.POINTER(Grid)]
    L.vox_solid_bits.argtypes = [P, C.POINTER(Grid), P, P, P]
    L.vox_voxelize_nodes.argtypes = [P, C.POINTER(Grid), I32, I32, P, P, P]
    L.vox_voxelize_nodes_noalloc.argtypes = [P, C.POINTER(Grid), I32, I32, P, P, P]
    L.vox_voxelize_partition.argtypes = [P, C.POINTER(Grid), I32, I32, P, P, P, P, P]
    L.vox_solid_bits_partition.argtypes = [P, C.POINTER(Grid), P, P, P, P]
    L.vox_voxelize_slice.argtypes = [P, I32, D, I32, I32, I32, U32, P, P, P, P]
    L.vox_node_dims.argtypes = [C.POINTER(Grid), I32, P]
    L.vox_node_dims.restype = None
    L.vox_node_count.argtypes = [C.POINTER(Grid), I32]
    L.vox_node_count.restype = C.c_size_t
    L.vox_node_size.argtypes = [I32]
    L.vox_node_size.restype = C.c_size_t
    L.vox_enable_timing.argtypes = [P, I32]
    L.vox_get_stage_times.argtypes = [P, C.POINTER(StageTimes)]
    F = C.c_float
    L.vox_set_schedule.argtypes = [P, I32]
    L.vox_set_heavy_threshold.argtypes = [P, F]
    L.vox_set_screen_ratio.argtypes = [P, F]
    L.vox_set_split_parts.argtypes = [P, U32]
    L.vox_cost_set_params.argtypes = [P, P]
    L.vox_cost_get_params.argtypes = [P, P]
    L.vox_cost_record.argtypes = [P, I32]
    L.vox_cost_samples.argtypes = [P, P, P, U32, C.POINTER(U32)]
    L.vox_cost_train.argtypes = [P, P, P, U32, U32, U32, F, U32, C.POINTER(F)]
    L.vox_mlp_forward.argtypes = [P, P, P, U32, P]
    L.vox_mlp_grad.argtypes = [P, P, P, P, U32, P, C.POINTER(F)]
    L.vox_voxelize_batch.argtypes = [P, C.POINTER(BatchItem), U32, I32, U32]
    L.vox_graph_create.argtypes = [P, C.POINTER(BatchItem), U32, I32, C.POINTER(P)]
    L.vox_graph_launch.argtypes = [P]
    L.vox_graph_fallbacks.argtypes = [P]
    L.vox_graph_fallbacks.restype = U32
    L.vox_graph_destroy.argtypes = [P]
    L.vox_graph_destroy.restype = None
    L.vox_mesh_load.argtypes = [C.c_char_p, C.POINTER(MeshC), C.c_char_p, C.c_size_t]
    L.vox_mesh_free.argtypes = [C.POINTER(MeshC)]
    L.vox_mesh_free.restype = None
    L.vox_write_nodes_vtk.argtypes = [C.c_char_p, P, I32, P, P, D, I32]
    L.vox_write_bits_vtk.argtypes = [C.c_char_p, P, P, P, D, I32]
    L.vox_set_brick_dims.argtypes = [P, I32, P]
    L.vox_get_brick_dims.argtypes = [P, I32, P]
    L.vox_autotune_bricks.argtypes = [P, C.POINTER(Grid), I32, P, P, P, U32, P, C.POINTER(F)]
    _lib = L
    return L


def device_count() -> int:
    return int(lib().vox_device_count())


def plan_grid_host(vertices, mode: int, param: float, vox_split_x: int = 1024) -> Grid:
    """Grid plan computed on the host (no device); same result as
    Context.plan_grid for the same vertices."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    g = Grid()
    st = lib().vox_plan_grid_host(_ptr(v), v.size // 3, mode, float(param), vox_split_x,
                                  C.byref(g))
    if st != 0:
        raise VoxError(st, "vox_plan_grid_host rejected its arguments")
    return g


def load_mesh(path: str):
    """(vertices float32 [nv, 3], indices uint32 [nt, 3]) of an OBJ, PLY, STL
    or legacy VTK file (vox_mesh_load)."""
    m = MeshC()
    err = C.create_string_buffer(512)
    st = lib().vox_mesh_load(os.fsencode(path), C.byref(m), err, len(err))
    if st != 0:
        raise VoxError(st, err.value.decode(errors="replace"))
    try:
        v = np.ctypeslib.as_array(m.vertices, shape=(3 * m.nv,)).copy().reshape(-1, 3) \
            if m.nv else np.zeros((0, 3), np.float32)
        t = np.ctypeslib.as_array(m.indices, shape=(3 * m.nt,)).copy().reshape(-1, 3) \
            if m.nt else np.zeros((0, 3), np.uint32)
    finally:
        lib().vox_mesh_free(C.byref(m))
    return v.astype(np.float32, copy=False), t.astype(np.uint32, copy=False)


def write_nodes_vtk(path: str, nodes: np.ndarray, kind: int, dims, origin=(0.0, 0.0, 0.0),
                    spacing: float = 1.0, binary: bool = True):
    """A node array as a legacy VTK STRUCTURED_POINTS file (vox_write_nodes_vtk)."""
    nodes = np.ascontiguousarray(nodes)
    n = int(dims[0]) * int(dims[1]) * int(dims[2])
    if nodes.nbytes != n * NODE_SIZE[kind]:
        raise ValueError("node array size does not match dims")
    st = lib().vox_write_nodes_vtk(os.fsencode(path), _ptr(nodes), kind,
                                   _ptr(np.array(dims, np.uint32)),
                                   _ptr(np.array(origin, np.float64)), float(spacing),
                                   1 if binary else 0)
    if st != 0:
        raise VoxError(st, f"cannot write {path}")


def write_bits_vtk(path: str, bits: np.ndarray, dims, origin=(0.0, 0.0, 0.0),
                   spacing: float = 1.0, binary: bool = True):
    """A solid bit array as a legacy VTK STRUCTURED_POINTS file (vox_write_bits_vtk)."""
    bits = np.ascontiguousarray(bits, np.uint32)
    if dims[0] % 32 or bits.size != dims[0] // 32 * dims[1] * dims[2]:
        raise ValueError("bit array size does not match dims")
    st = lib().vox_write_bits_vtk(os.fsencode(path), _ptr(bits), _ptr(np.array(dims, np.uint32)),
                                  _ptr(np.array(origin, np.float64)), float(spacing),
                                  1 if binary else 0)
    if st != 0:
        raise VoxError(st, f"cannot write {path}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Graph:
    """A captured fused batch: launch() replays it (vox_graph_launch)."""

    def __init__(self, ctx: "Context", h, items):
        self.ctx, self.h, self.items = ctx, h, items

    def launch(self):
        st = lib().vox_graph_launch(self.h)
        if st != OK:
            raise VoxError(st, "vox_graph_launch failed")

    @property
    def fallbacks(self) -> int:
        return int(lib().vox_graph_fallbacks(self.h))

    def close(self):
        if self.h:
            lib().vox_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device allocation owned by a Context (vox_alloc / vox_free)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        ctx._check(lib().vox_alloc(ctx.h, max(1, self.nbytes), C.byref(p)))
        self.ptr = p

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx._check(lib().vox_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), a.nbytes))

    def download(self, dtype, count=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dt)
        self.ctx._check(lib().vox_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ctx.h:
            lib().vox_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One vox_ctx: a device, a stream and a device arena."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().vox_create(device, C.byref(h))
        if st != OK:
            raise VoxError(st, "vox_create failed (no HIP device?)")
        self.h = h
        self._keep = []

    def close(self):
        if getattr(self, "h", None):
            lib().vox_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int):
        if st != OK:
            msg = lib().vox_last_error(self.h)
            raise VoxError(st, msg.decode() if msg else "")

    # -- mesh ---------------------------------------------------------------
    def set_mesh(self, vertices: np.ndarray, indices: np.ndarray):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        self._check(lib().vox_set_mesh(self.h, _ptr(v), v.shape[0], _ptr(t), t.shape[0]))
        self.nt = t.shape[0]

    def set_mesh_device(self, vptr: int, nv: int, iptr: int, nt: int):
        self._check(lib().vox_set_mesh_device(self.h, C.c_void_p(vptr), nv, C.c_void_p(iptr), nt))
        self.nt = nt

    def set_materials(self, materials, n_unique: int = 0):
        if materials is None:
            self._check(lib().vox_set_materials(self.h, None, 0))
            return
        m = np.ascontiguousarray(materials, np.uint8)
        self._check(lib().vox_set_materials(self.h, _ptr(m), n_unique))

    def set_materials_device(self, mptr: int, n_unique: int):
        self._check(lib().vox_set_materials_device(self.h, C.c_void_p(mptr), n_unique))

    # -- plan ---------------------------------------------------------------
    def plan_grid(self, mode: int, param: float, vox_split_x: int = 1024) -> Grid:
        g = Grid()
        self._check(lib().vox_plan_grid(self.h, mode, float(param), vox_split_x, C.byref(g)))
        return g

    # -- stages -------------------------------------------------------------
    def solid_bits_device(self, grid: Grid, out_ptr, shift=None, split_yz=(512, 512)):
        sh = np.array(shift if shift is not None else (0.0, 0.0, 0.0), np.float64)
        sp = np.array(split_yz, np.uint32)
        self._check(lib().vox_solid_bits(self.h, C.byref(grid), _ptr(sh), _ptr(sp),
                                         C.c_void_p(int(out_ptr) if not isinstance(out_ptr, C.c_void_p) else out_ptr.value)))

    def solid_bits(self, grid: Grid, shift=None, split_yz=(512, 512)) -> np.ndarray:
        words = grid.alloc[0] // 32 * grid.alloc[1] * grid.alloc[2]
        buf = DeviceBuffer(self, words * 4)
        self.solid_bits_device(grid, buf.ptr.value, shift, split_yz)
        out = buf.download(np.uint32, words)
        buf.free()
        return out

    def node_count(self, grid: Grid, kind: int) -> int:
        return int(lib().vox_node_count(C.byref(grid), kind))

    def node_dims(self, grid: Grid, kind: int):
        d = np.zeros(3, np.uint32)
        lib().vox_node_dims(C.byref(grid), kind, _ptr(d))
        return tuple(int(x) for x in d)

    def voxelize_nodes_device(self, grid: Grid, kind: int, out_ptr: int, use_materials=True,
                              vox_split=DEFAULT_VOX_SPLIT, mat_split=DEFAULT_MAT_SPLIT):
        vs = np.array(vox_split, np.uint32)
        ms = np.array(mat_split, np.uint32)
        self._check(lib().vox_voxelize_nodes(self.h, C.byref(grid), kind, 1 if use_materials else 0,
                                             _ptr(vs), _ptr(ms), C.c_void_p(int(out_ptr))))

    def voxelize_nodes_noalloc_device(self, grid: Grid, kind: int, out_ptr: int,
                                      use_materials=True, vox_split=DEFAULT_VOX_SPLIT,
                                      mat_split=DEFAULT_MAT_SPLIT):
        """vox_voxelize_nodes_noalloc: the node pipeline with every scratch
        buffer in the library's per-device scratch array (no allocation)."""
        vs = np.array(vox_split, np.uint32)
        ms = np.array(mat_split, np.uint32)
        self._check(lib().vox_voxelize_nodes_noalloc(self.h, C.byref(grid), kind,
                                                     1 if use_materials else 0, _ptr(vs),
                                                     _ptr(ms), C.c_void_p(int(out_ptr))))

    def voxelize_partition(self, grid: Grid, kind: int, part_min, part_max, use_materials=True,
                           vox_split=DEFAULT_VOX_SPLIT, mat_split=DEFAULT_MAT_SPLIT) -> np.ndarray:
        """Nodes of the partition [part_min, part_max) of the grid's ext (y, z)
        space: res[0] x (py + 2) x (pz + 2) nodes (vox_voxelize_partition)."""
        py, pz = part_max[0] - part_min[0], part_max[1] - part_min[1]
        nbytes = grid.alloc[0] * (py + 2) * (pz + 2) * NODE_SIZE[kind]
        buf = DeviceBuffer(self, nbytes)
        self._check(lib().vox_voxelize_partition(
            self.h, C.byref(grid), kind, 1 if use_materials else 0,
            _ptr(np.array(vox_split, np.uint32)), _ptr(np.array(mat_split, np.uint32)),
            _ptr(np.array(part_min, np.uint32)), _ptr(np.array(part_max, np.uint32)), buf.ptr))
        out = buf.download(np.uint8, nbytes)
        buf.free()
        return out

    def solid_bits_partition(self, grid: Grid, part_min, part_max,
                             split_yz=(512, 512)) -> np.ndarray:
        """Solid bits of the partition [part_min, part_max) of the grid's ext
        (y, z) space: alloc[0]/32 x (py + 2) x (pz + 2) words
        (vox_solid_bits_partition)."""
        py, pz = part_max[0] - part_min[0], part_max[1] - part_min[1]
        words = grid.alloc[0] // 32 * (py + 2) * (pz + 2)
        buf = DeviceBuffer(self, words * 4)
        self._check(lib().vox_solid_bits_partition(
            self.h, C.byref(grid), _ptr(np.array(split_yz, np.uint32)),
            _ptr(np.array(part_min, np.uint32)), _ptr(np.array(part_max, np.uint32)), buf.ptr))
        out = buf.download(np.uint32, words)
        buf.free()
        return out

    def slice_dims(self, mode: int, param: float, kind: int, direction: int, slice_: int = 1,
                   vox_split=(1024, 512), mat_split=(512, 512)):
        d = np.zeros(3, np.uint32)
        self._check(lib().vox_voxelize_slice(self.h, mode, float(param), kind, 0, direction,
                                             slice_, _ptr(np.array(vox_split, np.uint32)),
                                             _ptr(np.array(mat_split, np.uint32)), None, _ptr(d)))
        return tuple(int(x) for x in d)

    def voxelize_slice(self, mode: int, param: float, kind: int, direction: int, slice_: int,
                       use_materials=True, vox_split=(1024, 512), mat_split=(512, 512)):
        """(nodes, dims) of one slice (vox_voxelize_slice); dims (x, y, z)."""
        dims = self.slice_dims(mode, param, kind, direction, slice_, vox_split, mat_split)
        nbytes = dims[0] * dims[1] * dims[2] * NODE_SIZE[kind]
        buf = DeviceBuffer(self, nbytes)
        d = np.zeros(3, np.uint32)
        self._check(lib().vox_voxelize_slice(self.h, mode, float(param), kind,
                                             1 if use_materials else 0, direction, slice_,
                                             _ptr(np.array(vox_split, np.uint32)),
                                             _ptr(np.array(mat_split, np.uint32)), buf.ptr,
                                             _ptr(d)))
        out = buf.download(np.uint8, nbytes)
        buf.free()
        return out, dims

    def voxelize_nodes(self, grid: Grid, kind: int, use_materials=True,
                       vox_split=DEFAULT_VOX_SPLIT, mat_split=DEFAULT_MAT_SPLIT) -> np.ndarray:
        n = int(lib().vox_node_count(C.byref(grid), kind))
        nbytes = n * NODE_SIZE[kind]
        buf = DeviceBuffer(self, nbytes)
        self.voxelize_nodes_device(grid, kind, buf.ptr.value, use_materials, vox_split, mat_split)
        out = buf.download(np.uint8, nbytes)
        buf.free()
        return out

    def sync(self):
        self._check(lib().vox_sync(self.h))

    def enable_timing(self, on: bool = True):
        self._check(lib().vox_enable_timing(self.h, 1 if on else 0))

    def stage_times(self) -> dict:
        t = StageTimes()
        self._check(lib().vox_get_stage_times(self.h, C.byref(t)))
        return t.as_dict()

    # -- brick schedule and cost model --------------------------------------
    def set_schedule(self, schedule: int):
        self._check(lib().vox_set_schedule(self.h, schedule))

    def voxelize_batch(self, items, schedule: int = 0, streams: int = 1):
        """items: a ctypes array of BatchItem (device pointers filled in).
        streams: 0 fused (each stage one launch over all items), 1 one item
        after another, k > 1 k worker streams."""
        self._check(lib().vox_voxelize_batch(self.h, items, len(items), schedule, streams))

    def graph(self, items, schedule: int = 0) -> "Graph":
        """A replayable fused batch (vox_graph_create); items as for
        voxelize_batch, fixed for the graph's life."""
        h = C.c_void_p()
        self._check(lib().vox_graph_create(self.h, items, len(items), schedule, C.byref(h)))
        return Graph(self, h, items)

    def set_brick_dims(self, dims, fcc: bool = False):
        d = np.array(dims, np.uint32)
        self._check(lib().vox_set_brick_dims(self.h, 1 if fcc else 0, _ptr(d)))

    def brick_dims(self, fcc: bool = False):
        d = np.zeros(3, np.uint32)
        self._check(lib().vox_get_brick_dims(self.h, 1 if fcc else 0, _ptr(d)))
        return tuple(int(x) for x in d)

    def autotune_bricks(self, grid: Grid, kind: int, out_ptr: int, reps: int = 2,
                        vox_split=DEFAULT_VOX_SPLIT, mat_split=DEFAULT_MAT_SPLIT):
        """Picks the fastest surface brick shape for this grid and node kind;
        returns (dims, ms)."""
        vs = np.array(vox_split, np.uint32)
        ms = np.array(mat_split, np.uint32)
        best = np.zeros(3, np.uint32)
        t = C.c_float()
        self._check(lib().vox_autotune_bricks(self.h, C.byref(grid), kind, _ptr(vs), _ptr(ms),
                                              C.c_void_p(int(out_ptr)), reps, _ptr(best),
                                              C.byref(t)))
        return tuple(int(x) for x in best), float(t.value)

    def set_heavy_threshold(self, threshold: float):
        self._check(lib().vox_set_heavy_threshold(self.h, float(threshold)))

    def set_split_parts(self, parts: int):
        """Sub-bricks per heavy brick at most (1, 2, 4; 0 = automatic)."""
        self._check(lib().vox_set_split_parts(self.h, int(parts)))

    def set_screen_ratio(self, ratio: float):
        """Bricks with at least `ratio` times the mean pair count are scored
        by the cost model (learned schedule)."""
        self._check(lib().vox_set_screen_ratio(self.h, float(ratio)))

    def cost_record(self, on: bool = True):
        self._check(lib().vox_cost_record(self.h, 1 if on else 0))

    def cost_samples(self):
        """(features (n, 16) float32, cycles (n,) float32) recorded so far."""
        n = C.c_uint32()
        self._check(lib().vox_cost_samples(self.h, None, None, 0, C.byref(n)))
        f = np.zeros((n.value, COST_FEATURES), np.float32)
        c = np.zeros(n.value, np.float32)
        self._check(lib().vox_cost_samples(self.h, _ptr(f), _ptr(c), n.value, C.byref(n)))
        return f, c

    def cost_train(self, features, cycles, epochs=30, batch=256, lr=2e-3, seed=0) -> float:
        f = np.ascontiguousarray(features, np.float32)
        c = np.ascontiguousarray(cycles, np.float32)
        mse = C.c_float()
        self._check(lib().vox_cost_train(self.h, _ptr(f), _ptr(c), c.size, epochs, batch,
                                         float(lr), seed, C.byref(mse)))
        return float(mse.value)

    def cost_params(self) -> np.ndarray:
        p = np.zeros(COST_PARAMS, np.float32)
        self._check(lib().vox_cost_get_params(self.h, _ptr(p)))
        return p

    def set_cost_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == COST_PARAMS
        self._check(lib().vox_cost_set_params(self.h, _ptr(p)))

    def mlp_forward(self, params, x) -> np.ndarray:
        """MLP outputs for rows of x (n, 16), computed by the HIP kernel."""
        p = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32).reshape(-1, COST_FEATURES)
        n = x.shape[0]
        bp, bx, by = DeviceBuffer(self, p.nbytes), DeviceBuffer(self, x.nbytes), DeviceBuffer(self, 4 * n)
        bp.upload(p)
        bx.upload(x)
        self._check(lib().vox_mlp_forward(self.h, bp.ptr, bx.ptr, n, by.ptr))
        return by.download(np.float32, n)

    def mlp_grad(self, params, x, t):
        """(gradient of mean squared error, loss) from the HIP kernels."""
        p = np.ascontiguousarray(params, np.float32)
        x = np.ascontiguousarray(x, np.float32).reshape(-1, COST_FEATURES)
        t = np.ascontiguousarray(t, np.float32)
        n = x.shape[0]
        bp, bx, bt = DeviceBuffer(self, p.nbytes), DeviceBuffer(self, x.nbytes), DeviceBuffer(self, t.nbytes)
        bg = DeviceBuffer(self, p.nbytes)
        bp.upload(p)
        bx.upload(x)
        bt.upload(t)
        loss = C.c_float()
        self._check(lib().vox_mlp_grad(self.h, bp.ptr, bx.ptr, bt.ptr, n, bg.ptr, C.byref(loss)))
        return bg.download(np.float32, COST_PARAMS), float(loss.value)


__all__ = ["Context", "Grid", "BatchItem", "DeviceBuffer", "Graph", "VoxError", "lib",
           "device_count",
           "plan_grid_host",
           "NODE_SHORT", "NODE_LONG", "NODE_PARTIAL", "NODE_SHORT_FCC", "NODE_LONG_FCC",
           "PLAN_MAX_DIMENSION", "PLAN_CUBE_LENGTH", "NODE_SIZE", "EXPORTED",
           "SCHEDULE_PLAIN", "SCHEDULE_LEARNED", "COST_FEATURES", "COST_PARAMS",
           "load_mesh", "write_nodes_vtk", "write_bits_vtk"]
