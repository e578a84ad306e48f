"""Boundary-id table of the oracle (src/voxelizer.cu:2142-2236).

Pinned against the reference's documentation of the ids
(include/voxelizer.h:47-83), which names the solid neighbours of each id in
words, and against the closed form "an id exists iff every axis has a solid
neighbour".  Permutation bits: -x (Left) 1, +x (Right) 2, -y (In) 4, +y (Out) 8,
-z (Down) 16, +z (Up) 32 (src/voxelizer.cu:2128-2135).
"""
import itertools

BITS = {"Left": 1, "Right": 2, "In": 4, "Out": 8, "Down": 16, "Up": 32}

# include/voxelizer.h:50-83, one line per id.
DOC = {
    1: "Down, Left, In", 2: "Down, Right, In", 3: "Down, Left, Out", 4: "Down, Right, Out",
    5: "Up, Left, In", 6: "Up, Right, In", 7: "Up, Left, Out", 8: "Up, Right, Out",
    9: "Down, Left, Right, In", 10: "Down, Left, Right, Out", 11: "Down, Left, In, Out",
    12: "Down, Right, In, Out", 13: "Up, Left, Right, In", 14: "Up, Left, Right, Out",
    15: "Up, Left, In, Out", 16: "Up, Right, In, Out", 17: "Up, Down, Left, In",
    18: "Up, Down, Right, In", 19: "Up, Down, Left, Out", 20: "Up, Down, Right, Out",
    21: "Left, Right, In, Out, Down", 22: "Left, Right, Out, Down, Up",
    23: "Left, Right, In, Down, Up", 24: "Right, In, Out, Down, Up",
    25: "Left, In, Out, Down, Up", 26: "Left, Right, In, Out, Up",
    27: "Left, Right, In, Out, Down, Up",
}


def test_lut_matches_documented_ids(oracle):
    for bid, names in DOC.items():
        perm = sum(BITS[n.strip()] for n in names.split(","))
        assert oracle.orientation(perm) == bid, (bid, names, perm)


def test_lut_zero_iff_an_axis_lacks_neighbours(oracle):
    ids = []
    for perm in range(64):
        valid = bool(perm & 3) and bool(perm & 12) and bool(perm & 48)
        got = oracle.orientation(perm)
        assert (got != 0) == valid, perm
        if valid:
            ids.append(got)
    assert sorted(ids) == list(range(1, 28))


def test_lut_covers_every_valid_combination_once(oracle):
    combos = list(itertools.product([1, 2, 3], repeat=3))  # per axis: -, +, both
    perms = {a | (b << 2) | (c << 4) for a, b, c in combos}
    assert len(perms) == 27
    assert {oracle.orientation(p) for p in perms} == set(range(1, 28))
