"""Partitions (one device's share of the reference's multi-device mode) and
slices (voxelizeSlice along x, y and z) on the GPU, bit-compared with the
oracle's restatement (pinned on the CPU in tests/test_oracle_slices.py)."""
import sys

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(ctx, m):
    ctx.set_mesh(m.vertices, m.indices)
    ctx.set_materials(m.materials, int(m.materials.max()) + 1)


@pytest.mark.parametrize("kind", [va.NODE_SHORT, va.NODE_LONG])
def test_partitions_match_oracle(gpu_ctx, oracle, kind):
    m = W.shoebox_room(seed=6, n_boxes=20, subdiv=3)
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.02)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.02)
    ry, rz = g.res[1], g.res[2]
    parts = [((0, 0), (ry, rz)), ((0, 0), (ry // 2, rz // 2)), ((ry // 2, 0), (ry, rz // 2)),
             ((ry // 3, rz // 4), (2 * ry // 3, 3 * rz // 4)), ((0, rz - 1), (ry, rz)),
             ((5, 0), (6, rz))]
    for vs in [(1024, 512, 512), (1024, 17, 9)]:
        for pmin, pmax in parts:
            say(f"partition {pmin}..{pmax} kind {kind} vs {vs}")
            got = gpu_ctx.voxelize_partition(g, kind, pmin, pmax, True, vs)
            ref = oracle.voxelize_part(m, og, kind, pmin, pmax, True, vs)
            assert np.array_equal(got, ref), (pmin, pmax, int(np.count_nonzero(got != ref)))


def test_bit_partitions_match_oracle(gpu_ctx, oracle):
    m = W.shoebox_room(seed=6, n_boxes=20, subdiv=3)
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.02)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.02)
    ry, rz = g.res[1], g.res[2]
    parts = [((0, 0), (ry, rz)), ((0, 0), (ry // 2, rz // 2)), ((ry // 2, rz // 2), (ry, rz)),
             ((ry // 3, rz // 4), (2 * ry // 3, 3 * rz // 4)), ((5, 0), (6, rz))]
    for split in [(512, 512), (17, 9)]:
        for pmin, pmax in parts:
            say(f"bit partition {pmin}..{pmax} split {split}")
            got = gpu_ctx.solid_bits_partition(g, pmin, pmax, split)
            ref = oracle.solid_bits_part(m, og, pmin, pmax, split)
            assert np.array_equal(got, ref), (pmin, pmax, int(np.count_nonzero(got != ref)))


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_slices_match_oracle(gpu_ctx, oracle, direction):
    m = W.shoebox_room(seed=8, n_boxes=16, subdiv=3)
    load(gpu_ctx, m)
    mode, param = va.PLAN_CUBE_LENGTH, 0.025
    vs = (1024, 1024) if direction == 0 else (1024, 64)
    # around the first and last slice and past the grid (the room is 5 x 4 x 3
    # m: 201, 161 and 121 slices along x, y and z at 2.5 cm)
    picks = [0, 1, 2, 60, 119, 120, 121, 122, 160, 161, 162, 200, 201, 202, 240]
    for s in picks:
        say(f"slice direction {direction} slice {s}")
        got, gd = gpu_ctx.voxelize_slice(mode, param, va.NODE_LONG, direction, s, True, vs)
        ref, rd = oracle.voxelize_slice(m, mode, param, va.NODE_LONG, direction, s, True, vs)
        assert gd == rd
        if not np.array_equal(got, ref):
            nb = va.NODE_SIZE[va.NODE_LONG]
            bad = np.unique(np.flatnonzero(got != ref) // nb)
            for n in bad[:8]:
                x, y, z = n % gd[0], (n // gd[0]) % gd[1], n // (gd[0] * gd[1])
                say(f"  node ({x}, {y}, {z}) of {gd}: got {got[n * nb:(n + 1) * nb]} "
                    f"ref {ref[n * nb:(n + 1) * nb]}")
        assert np.array_equal(got, ref), (s, int(np.count_nonzero(got != ref)))


def test_slice_short_nodes_and_mat_split(gpu_ctx, oracle):
    m = W.icosphere(seed=3)
    load(gpu_ctx, m)
    for direction in (0, 1, 2):
        for s in (5, 40, 60):
            got, gd = gpu_ctx.voxelize_slice(va.PLAN_MAX_DIMENSION, 100, va.NODE_SHORT, direction, s,
                                             True, (1024, 1024), (33, 20))
            ref, rd = oracle.voxelize_slice(m, va.PLAN_MAX_DIMENSION, 100, va.NODE_SHORT,
                                            direction, s, True, (1024, 1024), (33, 20))
            assert gd == rd and np.array_equal(got, ref), (direction, s)


def test_x_slice_needs_whole_rotated_rows(gpu_ctx):
    m = W.icosphere(seed=3)
    load(gpu_ctx, m)
    with pytest.raises(va.VoxError) as e:
        gpu_ctx.voxelize_slice(va.PLAN_MAX_DIMENSION, 100, va.NODE_SHORT, 0, 10, True, (1024, 64))
    assert e.value.status == va.ERR_UNSUPPORTED
