"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NODE_SHORT, NODE_LONG, NODE_PARTIAL, NODE_SHORT_FCC, NODE_LONG_FCC = range(5)
NODE_SIZE = {NODE_SHORT: 1, NODE_LONG: 2, NODE_PARTIAL: 8, NODE_SHORT_FCC: 2, NODE_LONG_FCC: 4}


class Grid(C.Structure):
    """Mirror of vox_grid (include/voxelizer_amd.h)."""
    _fields_ = [("min_vertex", C.c_double * 3), ("max_vertex", C.c_double * 3),
                ("voxel_length", C.c_double), ("res", C.c_uint32 * 3),
                ("alloc", C.c_uint32 * 3), ("x_splits", C.c_uint32), ("pad_", C.c_uint32)]

    def as_dict(self):
        return {"min_vertex": list(self.min_vertex), "max_vertex": list(self.max_vertex),
                "voxel_length": self.voxel_length, "res": list(self.res),
                "alloc": list(self.alloc), "x_splits": self.x_splits}


def build(force: bool = False) -> str:
    srcs = [os.path.join(HERE, f) for f in ("oracle.cpp", "oracle.h")]
    stale = not os.path.exists(LIB_PATH) or any(
        os.path.getmtime(s) > os.path.getmtime(LIB_PATH) for s in srcs if os.path.exists(s))
    if force or stale:
        subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_orientation.restype = C.c_uint8
        L.orc_orientation.argtypes = [C.c_uint8]
        L.orc_plan_grid.argtypes = [P, C.c_uint32, C.c_int, C.c_double, C.c_uint32, C.POINTER(Grid)]
        L.orc_solid_bits.argtypes = [P, P, C.c_uint32, C.POINTER(Grid), P, C.c_uint32, C.c_uint32, P]
        L.orc_voxelize_nodes.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.POINTER(Grid), C.c_int,
                                         P, P, P, C.c_int]
        L.orc_surface_keys.argtypes = [P, P, C.c_uint32, C.POINTER(Grid), P, P, P]
        L.orc_classify.argtypes = [P, P, C.c_uint32, P, C.c_float, P]
        L.orc_center_crossing.argtypes = [P, P, C.c_double, C.c_uint32, C.c_uint32, P]
        L.orc_erode_nodes.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_int, P]
        L.orc_fcc_boundaries.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32, P]
        L.orc_tile_overlaps.restype = C.c_uint32
        L.orc_tile_overlaps.argtypes = [P, P, C.c_double, C.c_uint32, C.c_uint32]
        L.orc_fragile.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.POINTER(Grid), C.c_int, P, P,
                                  C.c_double, C.c_double, P, P]
        L.orc_voxelize_part.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.POINTER(Grid), C.c_int,
                                        P, P, P, P, P, C.c_int]
        L.orc_cut_volume.restype = C.c_float
        L.orc_cut_volume.argtypes = [P, P, P, P, C.c_float]
        L.orc_solid_bits_part.argtypes = [P, P, C.c_uint32, C.POINTER(Grid), C.c_uint32,
                                          C.c_uint32, P, P, P]
        L.orc_voxelize_slice.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_int, C.c_double,
                                         C.c_int, C.c_int, C.c_uint32, P, P, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def orientation(perm: int) -> int:
    return int(lib().orc_orientation(perm))


def plan_grid(vertices: np.ndarray, mode: int, param: float, vox_split_x: int = 1024) -> Grid:
    v = np.ascontiguousarray(vertices, np.float32)
    g = Grid()
    r = lib().orc_plan_grid(_p(v), v.shape[0], mode, float(param), vox_split_x, C.byref(g))
    if r != 0:
        raise ValueError("orc_plan_grid failed")
    return g


def solid_bits(mesh, grid: Grid, shift=(0.0, 0.0, 0.0), split_yz=(512, 512)) -> np.ndarray:
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    X, Y, Z = grid.alloc
    bits = np.zeros(X // 32 * Y * Z, np.uint32)
    em = np.array([grid.min_vertex[a] + shift[a] for a in range(3)], np.float64)
    r = lib().orc_solid_bits(_p(v), _p(t), t.shape[0], C.byref(grid), _p(em), split_yz[0],
                             split_yz[1], _p(bits))
    if r != 0:
        raise ValueError("orc_solid_bits failed")
    return bits


def node_dims(grid: Grid, kind: int):
    X, Y, Z = grid.alloc
    if kind in (NODE_SHORT_FCC, NODE_LONG_FCC):
        return (2 * X, Y, 2 * Z)
    return (X, Y, Z)


def voxelize_nodes(mesh, grid: Grid, kind: int, use_materials: bool = True,
                   vox_split=(1024, 512, 512), mat_split=(512, 512, 512),
                   n_threads: int = 0) -> np.ndarray:
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    dims = node_dims(grid, kind)
    n = dims[0] * dims[1] * dims[2]
    out = np.zeros(n * NODE_SIZE[kind], np.uint8)
    mats = None
    if use_materials and mesh.materials is not None:
        mats = np.ascontiguousarray(mesh.materials, np.uint8)
    vs = np.array(vox_split, np.uint32)
    ms = np.array(mat_split, np.uint32)
    r = lib().orc_voxelize_nodes(_p(v), v.shape[0], _p(t), t.shape[0],
                                 _p(mats) if mats is not None else None, C.byref(grid), kind,
                                 _p(vs), _p(ms), _p(out), n_threads)
    if r != 0:
        raise ValueError("orc_voxelize_nodes failed (kind %d)" % kind)
    return out


def surface_keys(mesh, grid: Grid, shift=(0.0, 0.0, 0.0), mat_split=(512, 512, 512)) -> np.ndarray:
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    X, Y, Z = grid.alloc
    keys = np.zeros(X * Y * Z, np.uint32)
    em = np.array([grid.min_vertex[a] + shift[a] for a in range(3)], np.float64)
    ms = np.array(mat_split, np.uint32)
    lib().orc_surface_keys(_p(v), _p(t), t.shape[0], C.byref(grid), _p(em), _p(ms), _p(keys))
    return keys


def classify(mesh, model_min, d) -> np.ndarray:
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    mm = np.array(model_min, np.float32)
    out = np.zeros(t.shape[0], np.uint32)
    lib().orc_classify(_p(v), _p(t), t.shape[0], _p(mm), float(d), _p(out))
    return out


def center_crossing(tri, ext_min, d, cy, cz):
    tr = np.ascontiguousarray(tri, np.float32).ravel()
    em = np.array(ext_min, np.float64)
    vx = np.zeros(1, np.uint32)
    ok = lib().orc_center_crossing(_p(tr), _p(em), float(d), cy, cz, _p(vx))
    return bool(ok), int(vx[0])


def tile_overlaps(tri, ext_min, d, ext_y, ext_z) -> int:
    tr = np.ascontiguousarray(tri, np.float32).ravel()
    em = np.array(ext_min, np.float64)
    return int(lib().orc_tile_overlaps(_p(tr), _p(em), float(d), ext_y, ext_z))


def erode_nodes(solid_zyx: np.ndarray, split_yz=(512, 512), literal=True) -> np.ndarray:
    """Boundary ids after erosion of a solid grid indexed [z, y, x]."""
    s = np.ascontiguousarray(solid_zyx, np.uint8)
    Z, Y, X = s.shape
    out = np.zeros_like(s)
    r = lib().orc_erode_nodes(_p(s), X, Y, Z, split_yz[0], split_yz[1], 1 if literal else 0, _p(out))
    if r:
        raise ValueError("orc_erode_nodes failed")
    return out


def fcc_boundaries(solid_zyx: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(solid_zyx, np.uint8)
    Z, Y, X = s.shape
    out = np.zeros(s.shape, np.uint16)
    lib().orc_fcc_boundaries(_p(s), X, Y, Z, _p(out))
    return out


def bits_to_grid(bits: np.ndarray, alloc) -> np.ndarray:
    """Unpack MSB-first bit words to a bool array indexed [z, y, x]."""
    X, Y, Z = alloc
    b = bits.view(np.uint32).reshape(Z, Y, X // 32)
    # MSB first: bit 31 is x%32 == 0.
    be = b.astype(">u4").view(np.uint8).reshape(Z, Y, X // 32, 4)
    return np.unpackbits(be, axis=3, bitorder="big").reshape(Z, Y, X).astype(bool)


FRAGILE_SOLID, FRAGILE_SURF, FRAGILE_MULTI, FRAGILE_RACE = 1, 2, 4, 8
FRAGILE_COUNTS = ["solid_tests", "solid_fragile_tests", "solid_fragile_crossings",
                  "surface_tests", "surface_fragile_tests", "multi_writer_voxels",
                  "material_race_voxels", "flagged_voxels"]


def fragile(mesh, grid: Grid, kind: int, use_materials: bool = True,
            vox_split=(1024, 512, 512), mat_split=(512, 512, 512), eps64=1e-12, eps32=1e-5):
    """(flags per node, counts) of orc_fragile: nodes whose value rests on a
    test within eps of its threshold (where the original build's FMA
    contraction could decide otherwise) or on the racy material store."""
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    dims = node_dims(grid, kind)
    flags = np.zeros(dims[0] * dims[1] * dims[2], np.uint8)
    mats = None
    if use_materials and mesh.materials is not None:
        mats = np.ascontiguousarray(mesh.materials, np.uint8)
    vs = np.array(vox_split, np.uint32)
    ms = np.array(mat_split, np.uint32)
    counts = np.zeros(8, np.uint64)
    r = lib().orc_fragile(_p(v), v.shape[0], _p(t), t.shape[0],
                          _p(mats) if mats is not None else None, C.byref(grid), kind, _p(vs),
                          _p(ms), float(eps64), float(eps32), _p(flags), _p(counts))
    if r:
        raise ValueError("orc_fragile failed")
    return flags, dict(zip(FRAGILE_COUNTS, (int(c) for c in counts)))


def voxelize_part(mesh, grid: Grid, kind: int, part_min, part_max, use_materials: bool = True,
                  vox_split=(1024, 512, 512), mat_split=(512, 512, 512), n_threads: int = 0):
    """Nodes of one partition [part_min, part_max) (y, z ext coordinates):
    res[0] x (py + 2) x (pz + 2) nodes (orc_voxelize_part)."""
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    py, pz = part_max[0] - part_min[0], part_max[1] - part_min[1]
    out = np.zeros(grid.alloc[0] * (py + 2) * (pz + 2) * NODE_SIZE[kind], np.uint8)
    mats = None
    if use_materials and mesh.materials is not None:
        mats = np.ascontiguousarray(mesh.materials, np.uint8)
    r = lib().orc_voxelize_part(_p(v), v.shape[0], _p(t), t.shape[0],
                                _p(mats) if mats is not None else None, C.byref(grid), kind,
                                _p(np.array(vox_split, np.uint32)),
                                _p(np.array(mat_split, np.uint32)),
                                _p(np.array(part_min, np.uint32)),
                                _p(np.array(part_max, np.uint32)), _p(out), n_threads)
    if r:
        raise ValueError("orc_voxelize_part failed (%d)" % r)
    return out


def cut_volume(tri, voxel, normal, m, d: float) -> float:
    """calculateVoxelPlaneIntersectionVolume as the oracle restates it
    (orc_cut_volume): tri (3, 3) float32, voxel int (3,), normal and m (3,)."""
    t = np.ascontiguousarray(tri, np.float32).reshape(9)
    return float(lib().orc_cut_volume(_p(t), _p(np.ascontiguousarray(voxel, np.int32)),
                                      _p(np.ascontiguousarray(normal, np.float32)),
                                      _p(np.ascontiguousarray(m, np.float32)), float(d)))


def solid_bits_part(mesh, grid: Grid, part_min, part_max, split_yz=(512, 512)) -> np.ndarray:
    """Solid bits of one partition [part_min, part_max) (y, z ext
    coordinates): alloc[0]/32 x (py + 2) x (pz + 2) words (orc_solid_bits_part)."""
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    py, pz = part_max[0] - part_min[0], part_max[1] - part_min[1]
    bits = np.zeros(grid.alloc[0] // 32 * (py + 2) * (pz + 2), np.uint32)
    r = lib().orc_solid_bits_part(_p(v), _p(t), t.shape[0], C.byref(grid), split_yz[0],
                                  split_yz[1], _p(np.array(part_min, np.uint32)),
                                  _p(np.array(part_max, np.uint32)), _p(bits))
    if r:
        raise ValueError("orc_solid_bits_part failed (%d)" % r)
    return bits


def voxelize_slice(mesh, mode: int, param: float, kind: int, direction: int, slice_: int,
                   use_materials: bool = True, vox_split=(1024, 512), mat_split=(512, 512)):
    """(nodes, dims) of one slice (orc_voxelize_slice); dims (x, y, z)."""
    v = np.ascontiguousarray(mesh.vertices, np.float32)
    t = np.ascontiguousarray(mesh.indices, np.uint32)
    mats = None
    if use_materials and mesh.materials is not None:
        mats = np.ascontiguousarray(mesh.materials, np.uint8)
    dims = np.zeros(3, np.uint32)
    vs = np.array(vox_split, np.uint32)
    ms = np.array(mat_split, np.uint32)
    args = (_p(v), v.shape[0], _p(t), t.shape[0], _p(mats) if mats is not None else None,
            mode, float(param), kind, direction, slice_, _p(vs), _p(ms))
    r = lib().orc_voxelize_slice(*args, None, _p(dims))
    if r:
        raise ValueError("orc_voxelize_slice failed (%d)" % r)
    out = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]) * NODE_SIZE[kind], np.uint8)
    lib().orc_voxelize_slice(*args, _p(out), _p(dims))
    return out, tuple(int(x) for x in dims)
