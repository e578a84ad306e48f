"""GPU parity: the HIP path against the CPU oracle, bit for bit, through the
C ABI.  Covers the workload configurations at their full sizes (icosphere
128^3, shoebox room 512x403x303, FCC room 640x255x384, large hall
1536x1024x768), every node type, materials on/off, y/z and material
subspaces, FCC shifts of the solid grid, open and degenerate meshes, tiles
and bricks holding from 1 to 130 triangles, and size-independent properties
at full size."""
import sys

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(ctx, m, n_unique=None):
    ctx.set_mesh(m.vertices, m.indices)
    if m.materials is not None:
        ctx.set_materials(m.materials, n_unique or int(m.materials.max()) + 1)


def diff_report(got, ref, size):
    bad = np.nonzero(got != ref)[0]
    return f"{np.unique(bad // size).size} nodes differ, first byte offsets {bad[:8].tolist()}"


def check_nodes(ctx, oracle, m, mode, param, kind, mats=True, vs=va.DEFAULT_VOX_SPLIT,
                ms=va.DEFAULT_MAT_SPLIT):
    say(f"case {m.name} nt={m.nt} mode={mode} param={param} kind={kind} mats={mats} vs={vs} ms={ms}")
    load(ctx, m)
    g = ctx.plan_grid(mode, param, vs[0])
    og = oracle.plan_grid(m.vertices, mode, param, vs[0])
    assert g.same(og), (g.as_dict(), og.as_dict())
    got = ctx.voxelize_nodes(g, kind, use_materials=mats, vox_split=vs, mat_split=ms)
    ref = oracle.voxelize_nodes(m, og, kind, mats, vox_split=vs, mat_split=ms)
    assert got.size == ref.size
    assert np.array_equal(got, ref), diff_report(got, ref, va.NODE_SIZE[kind])
    return got


@pytest.mark.parametrize("kind", [va.NODE_SHORT, va.NODE_LONG, va.NODE_SHORT_FCC, va.NODE_LONG_FCC])
@pytest.mark.parametrize("mats", [True, False])
def test_icosphere_all_kinds(gpu_ctx, oracle, kind, mats):
    m = W.icosphere(seed=0)
    if kind in (va.NODE_SHORT_FCC, va.NODE_LONG_FCC):
        check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.008, kind, mats)
    else:
        check_nodes(gpu_ctx, oracle, m, va.PLAN_MAX_DIMENSION, 126, kind, mats)


def test_room_long_full_size(gpu_ctx, oracle):
    m = W.shoebox_room()
    nodes = check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.01, va.NODE_LONG)
    assert nodes.size == 512 * 403 * 303 * 2


@pytest.mark.parametrize("kind", [va.NODE_LONG_FCC, va.NODE_SHORT_FCC])
def test_room_fcc_full_size(gpu_ctx, oracle, kind):
    m = W.shoebox_room()
    a = 4.0 ** (1.0 / 3.0) * 0.01
    nodes = check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, a, kind)
    assert nodes.size == 640 * 255 * 384 * va.NODE_SIZE[kind]


@pytest.mark.parametrize("vs,ms", [
    ((1024, 100, 64), (100, 150, 77)),
    ((1024, 33, 17), (512, 512, 512)),
    ((64, 512, 512), (33, 29, 41)),
    ((1024, 512, 1), (512, 512, 512)),
])
def test_room_subspaces(gpu_ctx, oracle, vs, ms):
    m = W.shoebox_room(seed=11, n_boxes=16, subdiv=3)
    check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.02, va.NODE_SHORT, True, vs, ms)


def test_fcc_subspaces(gpu_ctx, oracle):
    m = W.icosphere(seed=1)
    check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.01, va.NODE_SHORT_FCC, True,
                (1024, 30, 20), (40, 33, 50))


@pytest.mark.parametrize("seed", [1, 2])
def test_open_soup(gpu_ctx, oracle, seed):
    m = W.random_soup(seed, 400)
    check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.01, va.NODE_LONG)


def test_degenerate_and_tiny_triangles(gpu_ctx, oracle):
    rng = np.random.default_rng(5)
    base = W.icosphere(seed=2, subdiv=2)
    v = base.vertices.astype(np.float64)
    t = base.indices.astype(np.int64)
    extra_v = [v[0], v[0], v[1], v[2], v[2] + 1e-6, v[2] + 2e-6]  # repeated / collinear
    n0 = v.shape[0]
    b = W._Builder()
    b.add(np.concatenate([v, np.array(extra_v)]),
          np.concatenate([t, [[n0, n0 + 1, n0 + 2], [n0 + 3, n0 + 4, n0 + 5]]]))
    m = b.mesh("degenerate")
    m.materials = (1 + rng.integers(0, 7, m.nt)).astype(np.uint8)
    check_nodes(gpu_ctx, oracle, m, va.PLAN_MAX_DIMENSION, 64, va.NODE_SHORT)


def test_maxdim_plans(gpu_ctx, oracle):
    for R in (16, 37, 200):
        m = W.box((0.1, -0.3, 0.2), (0.7, 0.4, 1.1), subdiv=3)
        check_nodes(gpu_ctx, oracle, m, va.PLAN_MAX_DIMENSION, R, va.NODE_LONG)


@pytest.mark.parametrize("shift", [(0.0, 0.0, 0.0), (0.005, 0.005, 0.0), (0.0, 0.005, 0.005)])
def test_solid_bits_shift_and_splits(gpu_ctx, oracle, shift):
    m = W.shoebox_room(seed=2, n_boxes=10, subdiv=2)
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.01)
    og = oracle.plan_grid(m.vertices, 1, 0.01)
    for sp in [(512, 512), (97, 61)]:
        say(f"solid bits shift={shift} split={sp}")
        got = gpu_ctx.solid_bits(g, shift=shift, split_yz=sp)
        ref = oracle.solid_bits(m, og, shift=shift, split_yz=sp)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", [va.NODE_LONG, va.NODE_LONG_FCC])
def test_material_lookup_path(oracle, monkeypatch, kind):
    """Meshes above the key-embedding limit look winner materials up in
    memory instead of carrying them in the winner key; the limit is lowered
    to 0 here so the room takes that path."""
    monkeypatch.setenv("VOX_EMBED_TRIS", "0")
    ctx = va.Context(0)
    try:
        m = W.shoebox_room(seed=3, n_boxes=24, subdiv=3)
        check_nodes(ctx, oracle, m, va.PLAN_CUBE_LENGTH, 0.012, kind)
    finally:
        ctx.close()


def stacked_tiles_mesh(d):
    """A unit box plus, per count n, a stack of n triangles in planes x = const
    over one 4x4 tile of voxel-centre rows (stack k covers centre rows
    8k+1..8k+3 in y and 5..7 in z), so tiles hold from 1 to 130 triangles:
    every split of a tile's rows over lanes, and several passes per wave."""
    counts = [1, 2, 3, 4, 5, 8, 9, 16, 17, 32, 33, 63, 64, 65, 130]
    rng = np.random.default_rng(11)
    b = W._Builder()
    box = W.box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), subdiv=3)
    b.add(box.vertices, box.indices, np.full(box.nt, 1, np.uint8))
    for k, n in enumerate(counts):
        y0, z0 = (8 * k + 0.3) * d, 4.3 * d
        xs = 0.05 + 0.9 * (np.arange(n) + rng.random(n) * 0.5) / n
        v = np.zeros((3 * n, 3))
        v[:, 0] = np.repeat(xs, 3)
        v[0::3, 1:] = (y0, z0)
        v[1::3, 1:] = (y0 + 6.8 * d, z0)
        v[2::3, 1:] = (y0, z0 + 6.8 * d)
        b.add(v, np.arange(3 * n).reshape(-1, 3), 1 + rng.integers(0, 7, n))
    return b.mesh("stacked_tiles")


@pytest.mark.parametrize("kind", [va.NODE_SHORT, va.NODE_LONG_FCC])
def test_stacked_tiles(gpu_ctx, oracle, kind):
    d = 0.0075
    m = stacked_tiles_mesh(d)
    check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, d, kind)


def test_repeatable(gpu_ctx):
    m = W.shoebox_room(seed=3, n_boxes=12, subdiv=3)
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.015)
    a = gpu_ctx.voxelize_nodes(g, va.NODE_LONG)
    b = gpu_ctx.voxelize_nodes(g, va.NODE_LONG)
    assert np.array_equal(a, b)


def test_hall_full_size(gpu_ctx, oracle):
    """The 1.2 G-voxel workload, bit-compared with the oracle (multi-core)."""
    m = W.hall()
    nodes = check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, 1.0 / 64, va.NODE_SHORT, True,
                        (1024, 1024, 768), (512, 512, 512))
    X, Y, Z = 1536, 1024, 768
    n = nodes.reshape(Z, Y, X)
    # size-independent properties: padding shell is empty, materials 0..7
    assert not n[0].any() and not n[-1].any() and not n[:, 0].any() and not n[:, -1].any()
    assert not n[:, :, 0].any()
    assert (n & 31).max() <= 27


@pytest.mark.parametrize("d,vs", [(0.5, (1024, 5, 7)), (0.25, (1024, 3, 2)),
                                  (0.25, (1024, 512, 512)), (0.25, (1024, 9, 1))])
def test_disagreeing_slivers_across_seams(gpu_ctx, oracle, d, vs):
    """Triangles whose solid-filter orientation (tN) and fill orientation
    (n.x) have different signs, lying across many y/z subspace seams: the
    per-subspace queue of the reference decides which tile rows count."""
    m = W.disagreeing_slivers()
    check_nodes(gpu_ctx, oracle, m, va.PLAN_CUBE_LENGTH, d, va.NODE_SHORT, True, vs)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, d)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, d)
    got = gpu_ctx.solid_bits(g, split_yz=vs[1:])
    assert np.array_equal(got, oracle.solid_bits(m, og, split_yz=vs[1:]))
