"""Grids wider than 32768 voxels in x (more than one 1024-word pass of the
solid fill's LDS rows), bit-compared with the oracle: crossing parity has to
carry from one pass into the next.  Three overlapping closed boxes along a
40 m bar at 1 mm voxels put crossings on both sides of the pass boundary
(x = 32768 voxels).  matSplitRes x is 1024 here: the surface pass takes at
most 64 matSplitRes subspaces per axis."""
import sys

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bar_mesh():
    parts = [W.box((0.0, 0.0, 0.0), (40.0, 0.012, 0.012), subdiv=3),
             W.box((20.3, 0.003, 0.003), (36.7, 0.009, 0.009), subdiv=2),
             W.box((32.5, 0.004, -0.001), (33.1, 0.0105, 0.006), subdiv=2)]
    v, t, m, base = [], [], [], 0
    for k, p in enumerate(parts):
        v.append(p.vertices)
        t.append(p.indices + base)
        m.append(np.full(p.nt, k + 2, np.uint8))
        base += p.nv
    return W.Mesh(np.concatenate(v).astype(np.float32), np.concatenate(t).astype(np.uint32),
                  np.concatenate(m), "bar")


@pytest.mark.parametrize("kind", [va.NODE_SHORT, va.NODE_LONG])
def test_wide_grid_matches_oracle(gpu_ctx, oracle, kind):
    m = bar_mesh()
    gpu_ctx.set_mesh(m.vertices, m.indices)
    gpu_ctx.set_materials(m.materials, int(m.materials.max()) + 1)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.001)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.001)
    assert g.same(og)
    assert g.alloc[0] > 32 * 1024, g.as_dict()
    say(f"wide grid {g.as_dict()['alloc']} kind {kind}")
    bits = gpu_ctx.solid_bits(g)
    assert np.array_equal(bits, oracle.solid_bits(m, og))
    ms = (1024, 512, 512)
    got = gpu_ctx.voxelize_nodes(g, kind, mat_split=ms)
    ref = oracle.voxelize_nodes(m, og, kind, True, mat_split=ms)
    assert np.array_equal(got, ref), np.flatnonzero(got != ref)[:8]
