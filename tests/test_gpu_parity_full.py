"""GPU parity at full size on the configurations the benchmark reports, each
against the CPU oracle bit for bit, through the C ABI:

  * the large hall (1536x1024x768 ShortNodes, ~1.4M triangles) exactly as
    bench.py times it: learned brick order (a cost model trained on a
    differently seeded hall), brick shape 128x8x4, one-item fused batch
    captured as a HIP graph and replayed;
  * the large hall at the reference's default voxSplitRes (1024, 512, 512):
    2 x 2 solid subspaces over the ext rows, 2 x 2 erosion subspaces, ids next
    to a later subspace from the pre-erosion grid;
  * the 256-mesh batch (icospheres, rooms, boxes; 500..200k triangles, 64..512
    voxels on the longest side, ShortNode or LongNode by material count)
    through the fused learned graph, every mesh compared with the oracle.
"""
import sys

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu

HALL_D = 1.0 / 64.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def train_model(ctx, mesh, mode, param, kind, vs, ms):
    """Per-brick durations recorded on `mesh`, model trained on them."""
    ctx.set_mesh(mesh.vertices, mesh.indices)
    ctx.set_materials(mesh.materials, int(mesh.materials.max()) + 1)
    g = ctx.plan_grid(mode, param)
    ctx.cost_record(True)
    ctx.voxelize_nodes(g, kind, True, vs, ms)
    f, c = ctx.cost_samples()
    ctx.cost_record(False)
    assert c.size > 100
    ctx.cost_train(f, c, epochs=20, batch=256, lr=2e-3, seed=0)


def device_items(ctx, meshes, kinds, grids, vs, ms, keep):
    items = (va.BatchItem * len(meshes))()
    outs = []
    for i, (m, kind, g) in enumerate(zip(meshes, kinds, grids)):
        arrs = (m.vertices, m.indices, m.materials)
        bufs = [va.DeviceBuffer(ctx, a.nbytes) for a in arrs]
        for b, a in zip(bufs, arrs):
            b.upload(np.ascontiguousarray(a))
        nb = ctx.node_count(g, kind) * va.NODE_SIZE[kind]
        out = va.DeviceBuffer(ctx, nb)
        keep.extend(bufs + [out])
        it = items[i]
        it.vertices_dev, it.nv = bufs[0].ptr.value, m.nv
        it.indices_dev, it.nt = bufs[1].ptr.value, m.nt
        it.materials_dev, it.n_unique = bufs[2].ptr.value, int(m.materials.max()) + 1
        it.kind, it.grid = kind, g
        it.vox_split[:] = vs
        it.mat_split[:] = ms
        it.nodes_dev = out.ptr.value
        outs.append((out, nb))
    return items, outs


def test_hall_learned_graph_as_benchmarked(oracle):
    vs, ms = (1024, 1024, 768), (512, 512, 512)
    ctx = va.Context(0)
    keep = []
    try:
        ctx.set_brick_dims((128, 8, 4))
        say("training the cost model on hall seed 5")
        train_model(ctx, W.hall(seed=5), va.PLAN_CUBE_LENGTH, HALL_D, va.NODE_SHORT, vs, ms)
        m = W.hall()
        g = va.plan_grid_host(m.vertices, va.PLAN_CUBE_LENGTH, HALL_D, vs[0])
        assert tuple(g.alloc) == (1536, 1024, 768)
        items, outs = device_items(ctx, [m], [va.NODE_SHORT], [g], vs, ms, keep)
        ctx.set_schedule(va.SCHEDULE_LEARNED)
        ctx.enable_timing(True)
        ctx.voxelize_batch(items, va.SCHEDULE_LEARNED, 0)
        assert ctx.stage_times()["learned_schedule"] == 1  # the model ordered the bricks
        ctx.enable_timing(False)
        graph = ctx.graph(items, va.SCHEDULE_LEARNED)
        say("hall: oracle")
        og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, HALL_D, vs[0])
        ref = oracle.voxelize_nodes(m, og, va.NODE_SHORT, True, vox_split=vs, mat_split=ms)
        out, nb = outs[0]
        try:
            for rep in range(2):
                out.upload(np.zeros(nb, np.uint8))
                graph.launch()
                say(f"hall: graph replay {rep}")
                got = out.download(np.uint8, nb)
                assert np.array_equal(got, ref), f"{int(np.count_nonzero(got != ref))} nodes differ"
            assert graph.fallbacks == 0
        finally:
            graph.close()
    finally:
        del keep[:]
        ctx.close()


def test_hall_default_vox_split(gpu_ctx, oracle):
    """voxSplitRes (1024, 512, 512): the solid filter per y/z subspace of the
    ext rows (1021 -> 2 parts, 765 -> 2 parts) and erosion per alloc subspace
    (1024 -> 2, 768 -> 2), as the reference's voxelizeWorker runs them."""
    vs, ms = (1024, 512, 512), (512, 512, 512)
    m = W.hall()
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(m.materials, 8)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, HALL_D, vs[0])
    say("hall default split: GPU")
    got = gpu_ctx.voxelize_nodes(g, va.NODE_SHORT, True, vs, ms)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, HALL_D, vs[0])
    assert g.same(og)
    say("hall default split: oracle")
    ref = oracle.voxelize_nodes(m, og, va.NODE_SHORT, True, vox_split=vs, mat_split=ms)
    assert np.array_equal(got, ref), f"{int(np.count_nonzero(got != ref))} nodes differ"


def test_batch_256_fused_learned_graph(oracle):
    vs, ms = (1024, 1024, 1024), (512, 512, 512)
    ctx = va.Context(0)
    keep = []
    try:
        say("training the cost model on room seed 5")
        train_model(ctx, W.shoebox_room(seed=5), va.PLAN_CUBE_LENGTH, 0.01, va.NODE_LONG,
                    (1024, 512, 512), ms)
        meshes = W.batch_meshes()
        assert len(meshes) == 256
        kinds = [va.NODE_SHORT if int(m.materials.max()) <= 7 else va.NODE_LONG for m in meshes]
        grids = [va.plan_grid_host(m.vertices, va.PLAN_CUBE_LENGTH, m.meta["cube_length"], vs[0])
                 for m in meshes]
        items, outs = device_items(ctx, meshes, kinds, grids, vs, ms, keep)
        graph = ctx.graph(items, va.SCHEDULE_LEARNED)
        try:
            for out, nb in outs:
                out.upload(np.zeros(nb, np.uint8))
            graph.launch()
            assert graph.fallbacks == 0
            total = 0
            for i, (m, kind, (out, nb)) in enumerate(zip(meshes, kinds, outs)):
                og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, m.meta["cube_length"],
                                      vs[0])
                assert grids[i].same(og), i
                ref = oracle.voxelize_nodes(m, og, kind, True, vox_split=vs, mat_split=ms)
                got = out.download(np.uint8, nb)
                assert np.array_equal(got, ref), (i, m.name, m.nt, tuple(og.alloc))
                total += nb // va.NODE_SIZE[kind]
                if i % 32 == 31:
                    say(f"batch: {i + 1} meshes bit-exact ({total} nodes)")
            assert total > 6e9
        finally:
            graph.close()
    finally:
        del keep[:]
        ctx.close()
