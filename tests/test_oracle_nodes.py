"""Node classification of the oracle.

Cartesian (src/voxelizer.cu:1791-1913 + host loop src/host_side.cpp:
1183-1206): pinned by a brute-force greatest fixed point computed here
("solid and a solid neighbour on each axis", removed until stable), the
literal in-place passes agreeing with snapshot passes, and the final state
being consistent (every id equals the table value of its final neighbours).

FCC (src/voxelizer.cu:1930-2048): pinned against neighbour solidity
recomputed from lattice positions, using the reference's documented lattice
(include/node_types.h:428-441): node m sits at (m_x, 2 m_y + (m_x+m_z) mod 2,
m_z) * a/2 and its 12 nearest neighbours are the offsets with two +-1
coordinates; bit j of the mask is the neighbour drawn at position j in the
diagram of include/node_types.h:408-424.
"""
import numpy as np
import pytest

from voxelizer_amd import workloads as W

LUT_BITS = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def neighbour_counts(s):
    """Per node: solid neighbour flags along -x,+x,-y,+y,-z,+z (zero padded)."""
    p = np.pad(s, 1)
    Z, Y, X = s.shape
    out = []
    for dx, dy, dz in LUT_BITS:
        out.append(p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X])
    return out


def brute_gfp(solid):
    s = solid.astype(bool).copy()
    while True:
        n = neighbour_counts(s)
        valid = s & (n[0] | n[1]) & (n[2] | n[3]) & (n[4] | n[5])
        if np.array_equal(valid, s):
            return s
        s = valid


def ids_of(oracle, s):
    n = neighbour_counts(s)
    perm = sum(f.astype(np.int32) << k for k, f in enumerate(n))
    lut = np.array([oracle.orientation(p) for p in range(64)], np.uint8)
    return np.where(s, lut[perm], 0).astype(np.uint8)


def random_solid(seed, shape, p):
    rng = np.random.default_rng(seed)
    s = rng.random(shape) < p
    # add some blobs so that big parts survive erosion
    Z, Y, X = shape
    s[Z // 4:3 * Z // 4, Y // 4:3 * Y // 4, X // 4:3 * X // 4] |= True
    return s


@pytest.mark.parametrize("seed,p", [(0, 0.3), (1, 0.6), (2, 0.85)])
def test_erosion_is_greatest_fixed_point(oracle, seed, p):
    s = random_solid(seed, (18, 21, 25), p)
    gfp = brute_gfp(s)
    for literal in (True, False):
        bid = oracle.erode_nodes(s, literal=literal)
        assert np.array_equal(bid > 0, gfp)
        assert np.array_equal(bid, ids_of(oracle, gfp))


def test_erosion_removed_nodes_were_invalid_and_idempotent(oracle):
    s = random_solid(5, (16, 16, 16), 0.5)
    bid = oracle.erode_nodes(s)
    again = oracle.erode_nodes(bid > 0)
    assert np.array_equal(again, bid)
    assert not np.any((bid > 0) & ~s)


def test_erosion_subspaces_keep_stale_ids(oracle):
    # With y/z subspaces the reference erodes them one after another and
    # never revisits earlier ones: ids there see later subspaces as they were
    # before their own erosion.  Restated here from that description.
    s = random_solid(9, (20, 20, 12), 0.55)
    Z, Y, X = s.shape
    split_z = 7  # -> subspaces z in [0,6), [6,13), [13,20) (floor division from the top)
    ref = s.copy()
    bounds = []
    prev = Z
    parts = (Z + split_z - 1) // split_z
    for i in range(parts, 0, -1):
        lo = prev - prev // i
        bounds.append((lo, prev))
        prev = lo
    bounds.reverse()
    ids = np.zeros(s.shape, np.uint8)
    for lo, hi in bounds:
        cur = ref.copy()
        while True:
            n = neighbour_counts(cur)
            valid = cur & (n[0] | n[1]) & (n[2] | n[3]) & (n[4] | n[5])
            nxt = cur.copy()
            nxt[lo:hi] = valid[lo:hi]
            if np.array_equal(nxt, cur):
                break
            cur = nxt
        ids[lo:hi] = ids_of(oracle, cur)[lo:hi]
        ref = cur
    got = oracle.erode_nodes(s, split_yz=(1024, split_z))
    assert np.array_equal(got, ids)
    assert not np.array_equal(got, oracle.erode_nodes(s))  # the seams do differ


def fcc_positions(shape):
    Z, Y, X = shape
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return np.stack([x, 2 * y + (x + z) % 2, z], -1)


FCC_OFFSETS = [(0, -1, -1), (-1, 0, -1), (1, 0, -1), (0, 1, -1),
               (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 1, 0),
               (0, -1, 1), (-1, 0, 1), (1, 0, 1), (0, 1, 1)]


@pytest.mark.parametrize("seed", [0, 1])
def test_fcc_mask_from_lattice_positions(oracle, seed):
    rng = np.random.default_rng(seed)
    shape = (10, 7, 12)
    s = rng.random(shape) < 0.6
    got = oracle.fcc_boundaries(s)
    pos = fcc_positions(shape)
    index = {tuple(p): i for i, p in enumerate(pos.reshape(-1, 3))}
    flat = s.ravel()
    exp = np.zeros(flat.size, np.uint16)
    for i, p in enumerate(pos.reshape(-1, 3)):
        if not flat[i]:
            continue
        m = 0
        for j, off in enumerate(FCC_OFFSETS):
            q = (p[0] + off[0], p[1] + off[1], p[2] + off[2])
            k = index.get(q)
            if k is not None and flat[k]:
                m |= 1 << j
        exp[i] = m
    assert np.array_equal(got.ravel(), exp)


def test_fcc_offsets_are_the_12_nearest(oracle):
    # Sanity of the restated lattice: all 12 offsets have length sqrt(2) in
    # units of a/2, i.e. a/sqrt(2), the FCC nearest-neighbour distance.
    for off in FCC_OFFSETS:
        assert sum(c * c for c in off) == 2
    assert len(set(FCC_OFFSETS)) == 12


def test_cartesian_output_is_self_consistent(oracle):
    m = W.shoebox_room(seed=3, n_boxes=6, subdiv=2)
    g = oracle.plan_grid(m.vertices, 1, 0.04)
    nodes = oracle.voxelize_nodes(m, g, oracle.NODE_LONG, True).reshape(-1, 2)
    X, Y, Z = g.alloc
    bid = nodes[:, 0].reshape(Z, Y, X)
    mat = nodes[:, 1].reshape(Z, Y, X)
    assert np.array_equal(bid, ids_of(oracle, bid > 0))
    # materials only on surface voxels, all from the mesh's ids (1..8)
    assert set(np.unique(mat)) <= set(range(0, 9))
    assert (mat > 0).sum() > 0


def test_short_node_packing(oracle):
    m = W.icosphere(seed=0, subdiv=2)
    m.materials = np.full(m.nt, 5, np.uint8)
    g = oracle.plan_grid(m.vertices, 0, 40)
    short = oracle.voxelize_nodes(m, g, oracle.NODE_SHORT, True)
    long_ = oracle.voxelize_nodes(m, g, oracle.NODE_LONG, True).reshape(-1, 2)
    assert np.array_equal(short, ((long_[:, 1].astype(np.uint16) << 5) | (long_[:, 0] & 31)).astype(np.uint8))
    # material ids of 8 or more wrap in the 3 bits of a ShortNode
    m.materials = np.full(m.nt, 13, np.uint8)
    short = oracle.voxelize_nodes(m, g, oracle.NODE_SHORT, True)
    assert set(np.unique(short >> 5)) <= {0, 13 & 7}


def ray_parity_at(mesh, pts, eps=1e-9):
    """(solid, certain) of arbitrary points by +x ray parity in float64: a
    point is inside when an odd number of triangles cross the ray from -inf
    to it; points whose y/z lies within eps of a triangle edge, or whose x
    lies within eps of a crossing, are uncertain."""
    V = mesh.vertices.astype(np.float64)
    P = np.asarray(pts, np.float64)
    count = np.zeros(len(P), np.int64)
    unsure = np.zeros(len(P), bool)
    scale = eps * (1.0 + np.abs(V).max()) ** 2
    for tri in mesh.indices:
        a, b, c = V[tri[0]], V[tri[1]], V[tri[2]]
        n = np.cross(b - a, c - a)
        lo = np.minimum(np.minimum(a, b), c) - 1e-6
        hi = np.maximum(np.maximum(a, b), c) + 1e-6
        near = (P[:, 1] >= lo[1]) & (P[:, 1] <= hi[1]) & (P[:, 2] >= lo[2]) & (P[:, 2] <= hi[2])
        if not near.any() or abs(n[0]) < 1e-300:
            continue
        Q = P[near]
        e = np.stack([(q[1] - p[1]) * (Q[:, 2] - p[2]) - (q[2] - p[2]) * (Q[:, 1] - p[1])
                      for p, q in ((a, b), (b, c), (c, a))])
        inside = np.all(e > scale, 0) | np.all(e < -scale, 0)
        edge = np.any(np.abs(e) <= scale, 0) & ~(np.any(e > scale, 0) & np.any(e < -scale, 0))
        xi = a[0] - (n[1] * (Q[:, 1] - a[1]) + n[2] * (Q[:, 2] - a[2])) / n[0]
        idx = np.nonzero(near)[0]
        count[idx] += inside & (xi < Q[:, 0] - eps)
        unsure[idx] |= edge | (inside & (np.abs(xi - Q[:, 0]) <= eps))
    return count % 2 == 1, ~unsure


@pytest.mark.parametrize("kind_name,a", [("LONG_FCC", 0.043), ("SHORT_FCC", 0.037)])
def test_fcc_nodes_sit_on_the_documented_lattice(oracle, kind_name, a):
    """Node (fx, fy, fz) of the FCC array lies at the documented lattice point
    (include/node_types.h:428-441) m * a/2 with y doubled and shifted by
    ((m_x + m_z) mod 2) * a/2, taken from the grid's min corner less one
    padding cell: p = min + (fx a/2 - a, (fy - 1) a + ((fx + fz) mod 2) a/2,
    fz a/2 - a).  Its solidity is checked against float64 ray parity at p and
    its mask's population against the solid lattice neighbours at
    p + (two of +-a/2): this pins the placement of the four shifted
    sub-grids, so e.g. swapping sub-grids 3 and 4 fails."""
    kind = getattr(oracle, "NODE_" + kind_name)
    m = W.icosphere(seed=1, subdiv=3)
    g = oracle.plan_grid(m.vertices, 1, a)
    FX, FY, FZ = oracle.node_dims(g, kind)
    raw = oracle.voxelize_nodes(m, g, kind, False)
    if kind == oracle.NODE_LONG_FCC:
        bid = raw.reshape(-1, 4)[:, :2].copy().view(np.uint16).ravel()
    else:
        bid = raw.view(np.uint16) & 4095
    fz, fy, fx = np.meshgrid(np.arange(FZ), np.arange(FY), np.arange(FX), indexing="ij")
    fx, fy, fz = fx.ravel(), fy.ravel(), fz.ravel()
    mn = np.array(g.min_vertex)
    d = g.voxel_length
    pos = np.stack([mn[0] + fx * (d / 2) - d,
                    mn[1] + (fy - 1) * d + ((fx + fz) % 2) * (d / 2),
                    mn[2] + fz * (d / 2) - d], 1)
    solid, cert = ray_parity_at(m, pos)
    h = d / 2
    offs = [(sx * h, sy * h, 0) for sx in (-1, 1) for sy in (-1, 1)] + \
           [(sx * h, 0, sz * h) for sx in (-1, 1) for sz in (-1, 1)] + \
           [(0, sy * h, sz * h) for sy in (-1, 1) for sz in (-1, 1)]
    nsolid = np.zeros(len(pos), np.int64)
    ncert = np.ones(len(pos), bool)
    for o in offs:
        s2, c2 = ray_parity_at(m, pos + np.array(o))
        nsolid += s2
        ncert &= c2
    pop = np.array([bin(int(b)).count("1") for b in bid])
    sure = cert & ncert
    assert sure.mean() > 0.98
    assert np.all(bid[sure & ~solid] == 0)
    assert np.array_equal(pop[sure & solid], nsolid[sure & solid])
    assert (sure & solid & (nsolid == 12)).sum() > 1000  # interior nodes: mask 4095
    assert (sure & solid & (nsolid < 12)).sum() > 200    # surface nodes
