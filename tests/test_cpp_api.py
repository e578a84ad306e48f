"""The C++ class vox::Voxelizer<Node> (include/vox/voxelizer.h), driven by the
compiled program tests/cpp/voxelizer_api_check the way code written against
the reference class (include/voxelizer.h:704-868) drives it.  Host checks run
without a device; the GPU cases compare its outputs with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "voxelizer_api_check")


def exe() -> str:
    if not os.path.exists(EXE):
        raise FileNotFoundError(f"{EXE} is not built (make -C tests/cpp)")
    return EXE


def write_mesh(path, m, n_unique):
    mats = m.materials
    with open(path, "wb") as f:
        f.write(np.array([m.nv, m.nt, 1 if mats is not None else 0, n_unique], np.uint32).tobytes())
        f.write(np.ascontiguousarray(m.vertices, np.float32).tobytes())
        f.write(np.ascontiguousarray(m.indices, np.uint32).tobytes())
        if mats is not None:
            f.write(np.ascontiguousarray(mats, np.uint8).tobytes())


def run(tmp_path, m, n_unique, kind, plain, mode, param, mats, vs=(1024, 512, 512),
        ms=(512, 512, 512), to_ram=1):
    mp, op = str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")
    write_mesh(mp, m, n_unique)
    args = [exe(), "run", mp, str(kind), str(plain), str(mode), repr(float(param)), str(int(mats)),
            *map(str, vs), *map(str, ms), str(to_ram), op]
    print(" ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    return r, op


def read_out(op):
    raw = np.fromfile(op, np.uint8)
    dim = tuple(int(x) for x in raw[:12].view(np.uint32))
    return dim, raw[12:]


def test_host_checks():
    r = subprocess.run([exe(), "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host ok" in r.stdout
    # Unsupported requests print the reference's messages.
    assert "FCC Nodes can currently only be defined through the" in r.stdout
    assert "Please use a standard Node-type when producing a" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mode,param,to_ram", [
    (va.NODE_SHORT, va.PLAN_MAX_DIMENSION, 126, 1),
    (va.NODE_SHORT, va.PLAN_MAX_DIMENSION, 126, 0),
    (va.NODE_LONG, va.PLAN_CUBE_LENGTH, 0.01, 0),
    (va.NODE_SHORT_FCC, va.PLAN_CUBE_LENGTH, 0.01, 1),
    (va.NODE_LONG_FCC, va.PLAN_CUBE_LENGTH, 0.012, 0),
])
def test_nodes_match_oracle(tmp_path, oracle, kind, mode, param, to_ram):
    m = W.icosphere(seed=3)
    r, op = run(tmp_path, m, 2, kind, 0, mode, param, True, to_ram=to_ram)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, got = read_out(op)
    og = oracle.plan_grid(m.vertices, mode, param, 1024)
    assert dim == tuple(oracle.node_dims(og, kind))
    ref = oracle.voxelize_nodes(m, og, kind, True)
    assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_room_long_nodes_with_splits(tmp_path, oracle):
    m = W.shoebox_room(seed=11, n_boxes=16, subdiv=3)
    vs, ms = (1024, 100, 64), (100, 150, 77)
    r, op = run(tmp_path, m, 9, va.NODE_LONG, 0, va.PLAN_CUBE_LENGTH, 0.02, True, vs, ms, 1)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, got = read_out(op)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.02, vs[0])
    assert np.array_equal(got, oracle.voxelize_nodes(m, og, va.NODE_LONG, True, vox_split=vs,
                                                     mat_split=ms))


@pytest.mark.gpu
@pytest.mark.parametrize("to_ram", [1, 0])
def test_plain_bits_match_oracle(tmp_path, oracle, to_ram):
    m = W.icosphere(seed=4)
    r, op = run(tmp_path, m, 2, va.NODE_SHORT, 1, va.PLAN_MAX_DIMENSION, 200, False,
                vs=(1024, 64, 96), to_ram=to_ram)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, got = read_out(op)
    og = oracle.plan_grid(m.vertices, va.PLAN_MAX_DIMENSION, 200, 1024)
    assert dim == (og.alloc[0], og.alloc[1], og.alloc[2])
    ref = oracle.solid_bits(m, og, split_yz=(64, 96))
    assert np.array_equal(got.view(np.uint32), ref)


@pytest.mark.gpu
def test_materials_off_and_partial_nodes(tmp_path, oracle):
    m = W.icosphere(seed=5)
    r, op = run(tmp_path, m, 2, va.NODE_LONG, 0, va.PLAN_CUBE_LENGTH, 0.01, False)
    assert r.returncode == 0, r.stdout + r.stderr
    _, got = read_out(op)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.01, 1024)
    assert np.array_equal(got, oracle.voxelize_nodes(m, og, va.NODE_LONG, False))
    # Voxelizer<PartialNode>: 8-byte nodes with the volume ratio
    for to_ram in (1, 0):
        r, op = run(tmp_path, m, 2, va.NODE_PARTIAL, 0, va.PLAN_CUBE_LENGTH, 0.01, True,
                    to_ram=to_ram)
        assert r.returncode == 0, r.stdout + r.stderr
        dim, got = read_out(op)
        assert dim == tuple(og.alloc)
        assert np.array_equal(got, oracle.voxelize_nodes(m, og, va.NODE_PARTIAL, True))


def rotated_about_z(m):
    """The model turned a quarter about z as rotateVertices does
    (src/host_side.cpp:3261-3290), in numpy."""
    v = m.vertices.astype(np.float64)
    mn, mx = v.min(0), v.max(0)
    cx, cy = (mx[0] - mn[0]) / 2.0, (mx[1] - mn[1]) / 2.0
    r = m.vertices.copy()
    r[:, 0] = ((cy + cx) - v[:, 1]).astype(np.float32)
    r[:, 1] = ((cy - cx) + v[:, 0]).astype(np.float32)
    return r


@pytest.mark.parametrize("kind,mode,param,along_x", [
    (va.NODE_SHORT, va.PLAN_MAX_DIMENSION, 150, 0),
    (va.NODE_SHORT, va.PLAN_MAX_DIMENSION, 150, 1),
    (va.NODE_LONG, va.PLAN_CUBE_LENGTH, 0.013, 1),
    (va.NODE_LONG_FCC, va.PLAN_CUBE_LENGTH, 0.02, 1),
])
def test_array_dimensions_match_oracle_plan(tmp_path, oracle, kind, mode, param, along_x):
    m = W.shoebox_room(seed=2, n_boxes=5, subdiv=2)
    mp = str(tmp_path / "mesh.bin")
    write_mesh(mp, m, 1)
    r = subprocess.run([exe(), "dims", mp, str(kind), str(mode), repr(float(param)), "1024",
                        str(along_x)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = tuple(int(x) for x in r.stdout.split()[1:4])
    g = oracle.plan_grid(rotated_about_z(m) if along_x else m.vertices, mode, param, 1024)
    X, Y, Z = g.alloc
    exp = (Y, X, Z) if along_x else (X, Y, Z)
    if kind in (va.NODE_SHORT_FCC, va.NODE_LONG_FCC):
        exp = (2 * exp[0], exp[1], 2 * exp[2])
    assert got == exp


@pytest.mark.gpu
@pytest.mark.parametrize("plain,kind,dc", [(0, va.NODE_LONG, (2, 2)), (0, va.NODE_SHORT, (3, 1)),
                                           (0, va.NODE_PARTIAL, (1, 2)),
                                           (1, va.NODE_SHORT, (1, 3))])
def test_simulated_devconfig_partitions_match_oracle(tmp_path, oracle, plain, kind, dc):
    """voxelize / voxelizeToNodes with a devConfig inside simulateMultidevice:
    one array per partition (z outer, y inner), each placed at (0, i % x,
    i / x), bit-equal to the oracle's partition."""
    r = W.shoebox_room(seed=6, n_boxes=12, subdiv=3)
    m = W.Mesh(r.vertices, r.indices, (r.materials % 8).astype(np.uint8), r.name)  # ShortNode
    mp, op = str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")
    write_mesh(mp, m, int(m.materials.max()) + 1)
    vs, ms = (1024, 40, 30), (512, 512, 512)
    args = [exe(), "parts", mp, str(kind), str(plain), str(va.PLAN_CUBE_LENGTH), "0.02", "1",
            str(dc[0]), str(dc[1]), *map(str, vs), *map(str, ms), op]
    print(" ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.02, vs[0])
    X = og.alloc[0]

    def bounds(n, k):  # splitResolutionByNrOfParts along one axis
        b, rest = [None] * k, n
        for i in range(k, 0, -1):
            b[i - 1] = (rest - rest // i, rest)
            rest = b[i - 1][0]
        return b

    ys, zs = bounds(og.res[1], dc[0]), bounds(og.res[2], dc[1])
    raw = np.fromfile(op, np.uint8)
    n = int(raw[:4].view(np.uint32)[0])
    assert n == dc[0] * dc[1]
    off = 4
    for i in range(n):
        hdr = tuple(int(x) for x in raw[off:off + 24].view(np.uint32))
        off += 24
        y, z = i % dc[0], i // dc[0]
        pmin, pmax = (ys[y][0], zs[z][0]), (ys[y][1], zs[z][1])
        assert hdr[3:] == (0, y, z)
        assert hdr[:3] == (X, pmax[0] - pmin[0] + 2, pmax[1] - pmin[1] + 2)
        if plain:
            nb = X // 32 * hdr[1] * hdr[2] * 4
            ref = oracle.solid_bits_part(m, og, pmin, pmax, vs[1:]).view(np.uint8)
        else:
            nb = X * hdr[1] * hdr[2] * va.NODE_SIZE[kind]
            ref = oracle.voxelize_part(m, og, kind, pmin, pmax, True, vs, ms)
        got = raw[off:off + nb]
        off += nb
        assert np.array_equal(got, ref), (i, pmin, pmax)
    assert off == raw.size


@pytest.mark.gpu
@pytest.mark.parametrize("direction,slice_,to_ram", [(0, 7, 1), (1, 33, 0), (2, 20, 1), (2, 0, 0)])
def test_slices_match_oracle(tmp_path, oracle, direction, slice_, to_ram):
    m = W.shoebox_room(seed=8, n_boxes=10, subdiv=3)
    mp, op = str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")
    write_mesh(mp, m, int(m.materials.max()) + 1)
    vs, ms = (1024, 1024), (512, 512)
    args = [exe(), "slice", mp, str(va.NODE_LONG), str(va.PLAN_CUBE_LENGTH), "0.03", "1",
            str(direction), str(slice_), *map(str, vs), *map(str, ms), str(to_ram), op]
    print(" ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dim, got = read_out(op)
    ref, rd = oracle.voxelize_slice(m, va.PLAN_CUBE_LENGTH, 0.03, va.NODE_LONG, direction, slice_,
                                    True, vs, ms)
    assert dim == tuple(rd)
    assert np.array_equal(got, ref)
