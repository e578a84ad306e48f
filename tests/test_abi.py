"""The C ABI library loads, exports every symbol include/voxelizer_amd.h
declares, and its pure (no-device) entry points behave.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import voxelizer_amd as va

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "voxelizer_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vox_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_exported_list():
    assert declared_functions() == sorted(va.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = va.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", va.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (vox_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported


def test_node_dims_and_sizes():
    g = va.Grid()
    g.alloc[0], g.alloc[1], g.alloc[2] = 512, 403, 303
    L = va.lib()
    assert L.vox_node_count(C.byref(g), va.NODE_LONG) == 512 * 403 * 303
    assert L.vox_node_count(C.byref(g), va.NODE_LONG_FCC) == 1024 * 403 * 606
    assert [L.vox_node_size(k) for k in range(5)] == [1, 2, 8, 2, 4]


def test_create_without_device_reports_no_device():
    if va.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(va.VoxError) as e:
        va.Context(0)
    assert e.value.status == va.ERR_NO_DEVICE


def test_build_info():
    assert b"gfx950" in va.lib().vox_build_info()


@pytest.mark.parametrize("mode,param,vsx", [(0, 126, 1024), (1, 0.01, 1024), (1, 0.013, 64),
                                            (0, 1000, 128)])
def test_host_plan_matches_oracle(oracle, mode, param, vsx):
    from voxelizer_amd import workloads as W
    for m in (W.icosphere(seed=0), W.shoebox_room(seed=1, n_boxes=4, subdiv=1)):
        g = va.plan_grid_host(m.vertices, mode, param, vsx)
        og = oracle.plan_grid(m.vertices, mode, param, vsx)
        assert g.same(og), (g.as_dict(), og.as_dict())


def test_host_plan_rejects_bad_arguments():
    v = np.zeros((3, 3), np.float32)
    for mode, param, vsx in [(0, 1, 1024), (1, 0.0, 1024), (1, 0.1, 0), (7, 1.0, 1024)]:
        with pytest.raises(va.VoxError):
            va.plan_grid_host(v, mode, param, vsx)
