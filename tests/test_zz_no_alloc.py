"""The entry point that allocates nothing (vox_voxelize_nodes_noalloc)
against vox_voxelize_nodes, bit for bit, through the C ABI: two calls in a
row on one context, the second on a smaller grid (its scratch regions reuse
the array the first call left behind), and the refusal of a grid whose
scratch exceeds the array."""
import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def run_both(ctx, m, mode, param, kind, vs=va.DEFAULT_VOX_SPLIT):
    ctx.set_mesh(m.vertices, m.indices)
    ctx.set_materials(m.materials, int(m.materials.max()) + 1)
    g = ctx.plan_grid(mode, param, vs[0])
    ref = ctx.voxelize_nodes(g, kind, True, vs)
    nb = ctx.node_count(g, kind) * va.NODE_SIZE[kind]
    out = va.DeviceBuffer(ctx, nb)
    out.upload(np.full(nb, 0xA5, np.uint8))  # every byte must be written
    ctx.voxelize_nodes_noalloc_device(g, kind, out.ptr.value, True, vs)
    got = out.download(np.uint8, nb)
    out.free()
    return ref, got


def test_noalloc_matches_voxelize_nodes_twice_then_smaller():
    ctx = va.Context(0)
    try:
        # 512 x 403 x 303 LongNode room (about 80 MB of scratch), then the
        # 128^3 ShortNode icosphere, then the room on the FCC lattice.
        cases = [(W.shoebox_room(), va.PLAN_CUBE_LENGTH, 0.01, va.NODE_LONG, (1024, 512, 512)),
                 (W.icosphere(seed=0), va.PLAN_MAX_DIMENSION, 126, va.NODE_SHORT,
                  (1024, 64, 48)),
                 (W.shoebox_room(seed=3, n_boxes=12), va.PLAN_CUBE_LENGTH, 0.02,
                  va.NODE_SHORT_FCC, (1024, 512, 512))]
        for m, mode, param, kind, vs in cases:
            ref, got = run_both(ctx, m, mode, param, kind, vs)
            assert np.array_equal(ref, got), (m.name, int(np.count_nonzero(ref != got)))
    finally:
        ctx.close()


def test_noalloc_refuses_a_grid_beyond_its_scratch():
    ctx = va.Context(0)
    try:
        m = W.shoebox_room(seed=2, n_boxes=4, subdiv=2)
        ctx.set_mesh(m.vertices, m.indices)
        ctx.set_materials(m.materials, 9)
        g = ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.004)  # about 1280 x 1000 x 750
        nb = ctx.node_count(g, va.NODE_SHORT)
        out = va.DeviceBuffer(ctx, nb)
        with pytest.raises(va.VoxError) as e:
            ctx.voxelize_nodes_noalloc_device(g, va.NODE_SHORT, out.ptr.value, True)
        assert e.value.status == va.ERR_UNSUPPORTED
        assert "MiB" in str(e.value)
        out.free()
        # the context still works afterwards, both ways
        ref, got = run_both(ctx, W.icosphere(seed=1), va.PLAN_MAX_DIMENSION, 64, va.NODE_LONG)
        assert np.array_equal(ref, got)
    finally:
        ctx.close()
