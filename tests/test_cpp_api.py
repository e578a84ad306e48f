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
def test_materials_off_and_partial_node_error(tmp_path, oracle):
    m = W.icosphere(seed=5)
    r, op = run(tmp_path, m, 2, va.NODE_LONG, 0, va.PLAN_CUBE_LENGTH, 0.01, False)
    assert r.returncode == 0, r.stdout + r.stderr
    _, got = read_out(op)
    og = oracle.plan_grid(m.vertices, va.PLAN_CUBE_LENGTH, 0.01, 1024)
    assert np.array_equal(got, oracle.voxelize_nodes(m, og, va.NODE_LONG, False))
    # PartialNode is not built: the class throws vox::Exception.
    r, _ = run(tmp_path, m, 2, va.NODE_PARTIAL, 0, va.PLAN_CUBE_LENGTH, 0.01, True)
    assert r.returncode == 3 and "PartialNode" in r.stderr, r.stdout + r.stderr
