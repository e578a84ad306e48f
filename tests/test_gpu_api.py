"""Error behaviour of the C ABI on the GPU (status codes mirror the
reference's exceptions / empty results, src/host_side.cpp:164-243)."""
import ctypes as C

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_materials_required_when_requested(gpu_ctx):
    m = W.icosphere()
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(None)
    g = gpu_ctx.plan_grid(va.PLAN_MAX_DIMENSION, 32)
    with pytest.raises(va.VoxError) as e:
        gpu_ctx.voxelize_nodes(g, va.NODE_SHORT, use_materials=True)
    assert e.value.status == va.ERR_INVALID
    out = gpu_ctx.voxelize_nodes(g, va.NODE_SHORT, use_materials=False)
    assert out.size == 34 * 34 * 34 or out.size > 0


def test_bad_index_rejected(gpu_ctx):
    v = np.zeros((3, 3), np.float32)
    t = np.array([[0, 1, 5]], np.uint32)
    with pytest.raises(va.VoxError) as e:
        gpu_ctx.set_mesh(v, t)
    assert e.value.status == va.ERR_INVALID


def test_bad_plan_arguments(gpu_ctx):
    m = W.icosphere()
    gpu_ctx.set_mesh(m.vertices, m.indices)
    for mode, p in [(va.PLAN_MAX_DIMENSION, 1), (va.PLAN_CUBE_LENGTH, 0.0),
                    (va.PLAN_CUBE_LENGTH, -1.0)]:
        with pytest.raises(va.VoxError) as e:
            gpu_ctx.plan_grid(mode, p)
        assert e.value.status == va.ERR_INVALID


def test_wide_x_split_rejected(gpu_ctx):
    # voxSplitRes.x above 1024 makes the reference's word bitmask undefined.
    m = W.box((0, 0, 0), (30.0, 1.0, 1.0))
    gpu_ctx.set_mesh(m.vertices, m.indices)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.01, 4096)
    with pytest.raises(va.VoxError) as e:
        gpu_ctx.solid_bits(g)
    assert e.value.status == va.ERR_UNSUPPORTED


def test_empty_triangle_list(gpu_ctx, oracle):
    v = W.icosphere().vertices
    gpu_ctx.set_mesh(v, np.zeros((0, 3), np.uint32))
    g = gpu_ctx.plan_grid(va.PLAN_MAX_DIMENSION, 40)
    out = gpu_ctx.voxelize_nodes(g, va.NODE_LONG, use_materials=False)
    assert not out.any()


def test_stage_times_reported(gpu_ctx):
    m = W.shoebox_room(seed=5, n_boxes=4, subdiv=2)
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(m.materials, 9)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.02)
    gpu_ctx.enable_timing(True)
    gpu_ctx.voxelize_nodes(g, va.NODE_LONG)
    t = gpu_ctx.stage_times()
    gpu_ctx.enable_timing(False)
    assert t["total_ms"] > 0 and t["nodes_ms"] > 0 and t["erosion_passes"] >= 1


@pytest.mark.parametrize("kind,shapes", [
    (va.NODE_LONG, [(64, 4, 8), (128, 2, 8), (128, 4, 4), (256, 4, 4), (256, 2, 8), (128, 8, 4),
                    (64, 1, 4)]),
    (va.NODE_LONG_FCC, [(64, 4, 8), (128, 4, 4), (64, 8, 4), (128, 2, 4), (64, 2, 8),
                        (64, 2, 2), (256, 1, 1)]),
])
def test_brick_shapes_give_identical_nodes(gpu_ctx, oracle, kind, shapes):
    fcc = kind == va.NODE_LONG_FCC
    m = W.shoebox_room(seed=31, n_boxes=12, subdiv=3)
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(m.materials, 9)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.017)
    default = gpu_ctx.brick_dims(fcc)
    ref = oracle.voxelize_nodes(m, oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.017), kind,
                                True)
    try:
        for sh in shapes:
            gpu_ctx.set_brick_dims(sh, fcc)
            assert gpu_ctx.brick_dims(fcc) == sh
            assert np.array_equal(gpu_ctx.voxelize_nodes(g, kind), ref), sh
    finally:
        gpu_ctx.set_brick_dims(default, fcc)


def test_brick_shape_validation(gpu_ctx):
    for bad in [(32, 4, 8), (128, 3, 8), (128, 8, 8), (256, 4, 8), (128, 4, 2)]:
        with pytest.raises(va.VoxError):
            gpu_ctx.set_brick_dims(bad, False)
    with pytest.raises(va.VoxError):
        gpu_ctx.set_brick_dims((128, 8, 8), True)
    with pytest.raises(va.VoxError):  # fewer than 256 voxels
        gpu_ctx.set_brick_dims((64, 1, 2), True)


def test_autotune_bricks(gpu_ctx, oracle):
    m = W.shoebox_room(seed=32, n_boxes=8, subdiv=3)
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(m.materials, 9)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.02)
    n = gpu_ctx.node_count(g, va.NODE_LONG) * 2
    buf = va.DeviceBuffer(gpu_ctx, n)
    default = gpu_ctx.brick_dims(False)
    try:
        dims, ms = gpu_ctx.autotune_bricks(g, va.NODE_LONG, buf.ptr.value, reps=1)
        assert gpu_ctx.brick_dims(False) == dims and ms > 0
        ref = oracle.voxelize_nodes(m, oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.02),
                                    va.NODE_LONG, True)
        assert np.array_equal(buf.download(np.uint8, n), ref)
    finally:
        gpu_ctx.set_brick_dims(default, False)
        buf.free()


def test_candidate_count_past_2_pow_32_is_rejected(gpu_ctx):
    """600 triangles that each span the whole y/z extent of a 32 x 8192 x
    16382 bit grid give 600 * 2048 * 4095 (about 5.0e9) (triangle, tile)
    candidates, more than the 32-bit offsets hold: the scan flags the wrap and
    the call fails before any pair kernel reads the offsets."""
    d = 1.0
    ny, nz = 8190, 16380  # ext resolution in y and z
    y1, z1 = (ny - 1) * d, (nz - 1) * d
    n = 600
    xs = np.linspace(1.0, 20.0, n)
    v = np.zeros((3 * n, 3), np.float32)
    v[:, 0] = np.repeat(xs, 3)
    v[0::3, 1:] = (0.0, 0.0)
    v[1::3, 1:] = (y1, 0.0)
    v[2::3, 1:] = (y1, z1)
    t = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    gpu_ctx.set_mesh(v, t)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, d)
    assert tuple(g.alloc) == (32, ny + 2, nz + 2)
    tiles = ((ny + 3) // 4) * ((nz + 3) // 4)
    assert n * tiles >= 2 ** 32
    with pytest.raises(va.VoxError) as e:
        gpu_ctx.solid_bits(g)
    assert e.value.status == va.ERR_UNSUPPORTED
    assert "2^32" in str(e.value)
