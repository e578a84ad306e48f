"""The per-voxel list of nodes where the original CUDA build could differ
(oracle orc_fragile; generator and committed counts in tests/golden/).

Pins: the committed counts of the two smaller workloads are reproduced; the
material race is flagged exactly where hand-built triangles of one class
with different materials share voxels (and not where they share a
material); fragile solid decisions mark whole row suffixes, the part of the
row whose parity they decide."""
import json
import os

import numpy as np
import pytest

from voxelizer_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden():
    sys_path = os.path.join(HERE, "golden")
    import importlib.util
    spec = importlib.util.spec_from_file_location("fragile_voxels",
                                                  os.path.join(sys_path, "fragile_voxels.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(os.path.join(sys_path, "fragile_counts.json")) as f:
        return mod, json.load(f)


@pytest.mark.parametrize("name", ["icosphere_128_short", "room_512x403x303_long"])
def test_committed_counts_reproduce(oracle, name):
    mod, golden = load_golden()
    out, _ = mod.analyse(name)
    assert out == golden[name]


def test_material_race_needs_different_materials(oracle):
    # Two coplanar 2D triangles (plane z = 0.5) sharing their hypotenuse,
    # both overlapping the voxels along it; in a second mesh the same two
    # triangles carry one material.
    v = np.array([[0.1, 0.1, 0.5], [0.9, 0.1, 0.5], [0.1, 0.9, 0.5], [0.9, 0.9, 0.5],
                  [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], np.float32)
    t = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    g = oracle.plan_grid(v, 1, 0.05)
    for mats, race in [((1, 2), True), ((3, 3), False)]:
        m = W.Mesh(v, t, np.array(mats, np.uint8), "pair")
        flags, counts = oracle.fragile(m, g, oracle.NODE_LONG, True)
        multi = (flags & oracle.FRAGILE_MULTI) != 0
        assert multi.sum() > 10
        raced = (flags & oracle.FRAGILE_RACE) != 0
        if race:
            assert np.array_equal(raced, multi)
        else:
            assert not raced.any()
        assert counts["multi_writer_voxels"] == multi.sum()


def test_fragile_solid_marks_row_suffixes(oracle):
    m = W.shoebox_room()
    g = oracle.plan_grid(m.vertices, 1, 0.01)
    flags, counts = oracle.fragile(m, g, oracle.NODE_LONG, True)
    X, Y, Z = g.alloc
    solid = ((flags & oracle.FRAGILE_SOLID) != 0).reshape(Z, Y, X)
    assert counts["solid_fragile_tests"] > 0 and solid.any()
    # once a row is marked, it stays marked to its end
    first = np.where(solid.any(2), solid.argmax(2), X)
    for z, y in zip(*np.nonzero(first < X)):
        assert solid[z, y, first[z, y]:].all()
