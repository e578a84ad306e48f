"""Grid planning of the oracle (src/host_side.cpp:2462-2634, 717-738).

Pinned by hand-computed integer cases (the dims of the benchmark workloads
worked out by hand from the reference's dimension rules, and small boxes
whose d divides their sides) and by the rounding rule of the x axis."""
import numpy as np
import pytest

from voxelizer_amd import workloads as W


def box_vertices(lo, hi):
    return W.box(lo, hi).vertices


def test_shoebox_room_dims(oracle):
    g = oracle.plan_grid(W.shoebox_room().vertices, 1, 0.01)
    assert list(g.res) == [512, 401, 301]
    assert list(g.alloc) == [512, 403, 303]
    assert g.x_splits == 1 and g.voxel_length == 0.01


def test_hall_dims_exact(oracle):
    m = W.hall()
    assert tuple(m.vertices.min(0)) == (0.0, 0.0, 0.0)
    assert tuple(m.vertices.max(0)) == (1533 / 64, 1021 / 64, 765 / 64)
    g = oracle.plan_grid(m.vertices, 1, 1.0 / 64)
    assert list(g.alloc) == [1536, 1024, 768]
    assert g.x_splits == 2  # 1536 > voxSplitRes.x = 1024


def test_fcc_room_dims(oracle):
    a = 4.0 ** (1.0 / 3.0) * 0.01
    g = oracle.plan_grid(W.shoebox_room().vertices, 1, a)
    assert list(g.alloc) == [320, 255, 192]


def test_icosphere_maxdim(oracle):
    g = oracle.plan_grid(W.icosphere().vertices, 0, 126)
    assert list(g.alloc) == [128, 128, 128]


@pytest.mark.parametrize("hi,R,expect_res", [
    ((10.0, 5.0, 2.0), 11, (32, 6, 3)),     # x longest: d = 1, x = R + 2 -> 13 -> 32
    ((5.0, 10.0, 2.0), 11, (32, 11, 3)),    # y longest: y = R, x = 5+1+2 = 8 -> 32
    ((2.0, 5.0, 10.0), 11, (32, 6, 11)),    # z longest
    ((8.0, 2.0, 8.0), 9, (32, 3, 9)),       # x == z: the z branch (strict compares)
    ((64.0, 1.0, 1.0), 65, (96, 2, 2)),     # 67 -> next multiple of 32
])
def test_maxdim_cases(oracle, hi, R, expect_res):
    g = oracle.plan_grid(box_vertices((0, 0, 0), hi), 0, R)
    assert tuple(g.res) == expect_res


@pytest.mark.parametrize("hi,d,expect_res", [
    ((1.0, 1.0, 1.0), 0.25, (32, 5, 5)),
    ((3.0, 2.0, 1.0), 0.5, (32, 5, 3)),
    ((31.0, 1.0, 1.0), 1.0, (64, 2, 2)),    # 31+1+2 = 34 -> 64
    ((29.0, 1.0, 1.0), 1.0, (32, 2, 2)),    # 29+1+2 = 32 stays
])
def test_cube_length_cases(oracle, hi, d, expect_res):
    g = oracle.plan_grid(box_vertices((0, 0, 0), hi), 1, d)
    assert tuple(g.res) == expect_res
    assert list(g.alloc) == [expect_res[0], expect_res[1] + 2, expect_res[2] + 2]


def test_x_split_rounding(oracle):
    # 2000 + 1 + 2 = 2003 voxels; voxSplitRes.x = 1024 -> 2 splits -> multiple of 64.
    g = oracle.plan_grid(box_vertices((0, 0, 0), (2000.0, 1.0, 1.0)), 1, 1.0, vox_split_x=1024)
    assert g.x_splits == 2 and g.res[0] == 2048
    g = oracle.plan_grid(box_vertices((0, 0, 0), (2000.0, 1.0, 1.0)), 1, 1.0, vox_split_x=512)
    assert g.x_splits == 4 and g.res[0] % 128 == 0 and g.res[0] >= 2003


def test_bbox_is_exact_min_max(oracle):
    rng = np.random.default_rng(3)
    v = (rng.random((1000, 3)) * 7 - 3).astype(np.float32)
    g = oracle.plan_grid(v, 1, 0.1)
    assert tuple(g.min_vertex) == tuple(float(x) for x in v.min(0))
    assert tuple(g.max_vertex) == tuple(float(x) for x in v.max(0))
