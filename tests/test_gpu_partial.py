"""PartialNode on the GPU through the C ABI, bit-compared with the oracle
(pinned in tests/test_oracle_partial.py): the whole node array, ratio bits
included (NaN stored as 0x7FC00000), for a sphere, a room with materials
split, the unit cube of the hand-derived 1.5 ratios, 1D slivers, partitions,
slices along x, y and z, and a fused batch that mixes node kinds."""
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


def unit_cube():
    v = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float32)
    q = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    t = []
    for a, b, c, e in q:
        t += [(a, b, c), (a, c, e)]
    return W.Mesh(v, np.array(t, np.uint32), np.arange(12, dtype=np.uint8) % 3 + 1, "cube")


def slivers():
    v = np.array([[0.0, 0.0, 0.0], [1.0, 0.001, 0.0], [0.5, 0.0, 0.001],
                  [0.0, -0.2, -0.2], [1.0, 0.3, 0.3], [0.2, 0.3, -0.2],
                  [0.1, 0.5, 0.1], [0.1001, 0.5, 0.9], [0.1, 0.5003, 0.5]], np.float32)
    t = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.uint32)
    return W.Mesh(v, t, np.array([4, 2, 7], np.uint8), "slivers")


@pytest.mark.parametrize("name,mode,param,mats,vs,ms", [
    ("icosphere", va.PLAN_MAX_DIMENSION, 120, True, (1024, 512, 512), (512, 512, 512)),
    ("icosphere", va.PLAN_MAX_DIMENSION, 120, False, (1024, 512, 512), (512, 512, 512)),
    ("room", va.PLAN_CUBE_LENGTH, 0.02, True, (1024, 100, 64), (100, 150, 77)),
    ("cube", va.PLAN_CUBE_LENGTH, 0.1, True, (1024, 512, 512), (512, 512, 512)),
    ("slivers", va.PLAN_CUBE_LENGTH, 0.05, True, (1024, 512, 512), (512, 512, 512)),
])
def test_partial_nodes_match_oracle(gpu_ctx, oracle, name, mode, param, mats, vs, ms):
    m = {"icosphere": lambda: W.icosphere(seed=3), "room": lambda: W.shoebox_room(seed=12),
         "cube": unit_cube, "slivers": slivers}[name]()
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(mode, param, vs[0])
    og = oracle.plan_grid(m.vertices, mode, param, vs[0])
    assert g.same(og)
    say(f"partial {name} mats={mats}")
    got = gpu_ctx.voxelize_nodes(g, va.NODE_PARTIAL, mats, vs, ms)
    ref = oracle.voxelize_nodes(m, og, va.NODE_PARTIAL, mats, vox_split=vs, mat_split=ms)
    assert got.size == ref.size
    bad = np.flatnonzero((got != ref).reshape(-1, 8).any(1))
    assert bad.size == 0, (bad.size, bad[:5], got.reshape(-1, 8)[bad[:3]], ref.reshape(-1, 8)[bad[:3]])
    r = ref.reshape(-1, 8)[:, 4:].copy().view(np.float32).reshape(-1)
    if name == "cube":
        assert np.isclose(r, 1.5, atol=1e-5).sum() > 500
    if name == "slivers":
        assert (ref.reshape(-1, 8)[:, 4:].copy().view(np.uint32) == 0x7FC00000).any()


def test_partial_partitions_and_slices(gpu_ctx, oracle):
    m = W.shoebox_room(seed=6, n_boxes=10, subdiv=3)
    load(gpu_ctx, m)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.025)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.025)
    ry, rz = g.res[1], g.res[2]
    for pmin, pmax in [((0, 0), (ry // 2, rz)), ((ry // 3, rz // 4), (2 * ry // 3, 3 * rz // 4))]:
        say(f"partial partition {pmin}..{pmax}")
        got = gpu_ctx.voxelize_partition(g, va.NODE_PARTIAL, pmin, pmax, True)
        ref = oracle.voxelize_part(m, og, va.NODE_PARTIAL, pmin, pmax, True)
        assert np.array_equal(got, ref), (pmin, pmax)
    for direction, s in [(0, 40), (1, 33), (2, 20), (2, 0)]:
        vs = (1024, 1024) if direction == 0 else (1024, 64)
        say(f"partial slice {direction} {s}")
        got, gd = gpu_ctx.voxelize_slice(va.PLAN_CUBE_LENGTH, 0.025, va.NODE_PARTIAL, direction,
                                         s, True, vs)
        ref, rd = oracle.voxelize_slice(m, va.PLAN_CUBE_LENGTH, 0.025, va.NODE_PARTIAL,
                                        direction, s, True, vs)
        assert gd == rd and np.array_equal(got, ref), (direction, s)


def test_fused_batch_mixing_partial_nodes(oracle):
    ctx = va.Context(0)
    keep = []
    try:
        meshes = [W.icosphere(seed=1), W.shoebox_room(seed=2, n_boxes=6, subdiv=2), unit_cube(),
                  W.icosphere(seed=4)]
        kinds = [va.NODE_PARTIAL, va.NODE_LONG, va.NODE_PARTIAL, va.NODE_SHORT]
        params = [60, 0.05, 0.1, 80]
        modes = [va.PLAN_MAX_DIMENSION, va.PLAN_CUBE_LENGTH, va.PLAN_CUBE_LENGTH,
                 va.PLAN_MAX_DIMENSION]
        vs, ms = (1024, 1024, 1024), (512, 512, 512)
        items = (va.BatchItem * len(meshes))()
        outs = []
        for i, m in enumerate(meshes):
            arrs = (m.vertices, m.indices, m.materials)
            bufs = [va.DeviceBuffer(ctx, a.nbytes) for a in arrs]
            for b, a in zip(bufs, arrs):
                b.upload(np.ascontiguousarray(a))
            g = va.plan_grid_host(m.vertices, modes[i], params[i], vs[0])
            nb = ctx.node_count(g, kinds[i]) * va.NODE_SIZE[kinds[i]]
            out = va.DeviceBuffer(ctx, nb)
            keep.extend(bufs + [out])
            it = items[i]
            it.vertices_dev, it.nv = bufs[0].ptr.value, m.nv
            it.indices_dev, it.nt = bufs[1].ptr.value, m.nt
            it.materials_dev, it.n_unique = bufs[2].ptr.value, int(m.materials.max()) + 1
            it.kind, it.grid = kinds[i], g
            it.vox_split[:] = vs
            it.mat_split[:] = ms
            it.nodes_dev = out.ptr.value
            outs.append((out, nb, g))
        for schedule in (va.SCHEDULE_PLAIN, va.SCHEDULE_LEARNED):
            for out, nb, _ in outs:
                out.upload(np.zeros(nb, np.uint8))
            say(f"partial batch schedule {schedule}")
            ctx.voxelize_batch(items, schedule, 0)
            for i, (m, (out, nb, g)) in enumerate(zip(meshes, outs)):
                og = oracle.plan_grid(m.vertices, modes[i], params[i], vs[0])
                ref = oracle.voxelize_nodes(m, og, kinds[i], True, vox_split=vs, mat_split=ms)
                assert np.array_equal(out.download(np.uint8, nb), ref), (schedule, i)
    finally:
        del keep[:]
        ctx.close()
