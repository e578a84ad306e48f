"""Mesh files and output dumps through the C ABI (vox_mesh_load,
vox_write_nodes_vtk, vox_write_bits_vtk), host only.

The readers stand in for the VTK readers the reference's tester uses
(tests/src/main.cpp:359-404, loadModel 911-955).  VTK is not available
here, so the pins are the published file formats: meshes written here by
independent numpy writers (OBJ, PLY ASCII / little / big endian, STL ASCII /
binary, legacy VTK ASCII / binary / DataFile 5.1) read back to the same
float vertices and triangles; hand-written files with known answers (OBJ
negative and slashed indices, polygons split into fans, VTK cell sections
other than POLYGONS ignored); STL's merge of coincident corners checked
against numpy's own first-appearance dedup.  The writers are read back by a
small parser here."""
import struct

import numpy as np
import pytest

import voxelizer_amd as va
from voxelizer_amd import workloads as W


@pytest.fixture(scope="module")
def mesh():
    m = W.shoebox_room(seed=3, n_boxes=4, subdiv=2)
    return np.ascontiguousarray(m.vertices, np.float32), np.ascontiguousarray(m.indices, np.uint32)


def fmt(x):
    return repr(float(np.float32(x)))


def write_obj(path, v, t):
    with open(path, "w") as f:
        f.write("# written by the test\no room\n")
        for p in v:
            f.write("v %s %s %s\n" % tuple(fmt(c) for c in p))
        f.write("vn 0 0 1\n")
        for tri in t:
            f.write("f %d//1 %d//1 %d//1\n" % tuple(int(i) + 1 for i in tri))


def write_ply(path, v, t, form):
    head = ["ply", f"format {form} 1.0", "comment test", f"element vertex {len(v)}",
            "property float x", "property float y", "property float z", "property uchar red",
            f"element face {len(t)}", "property list uchar int vertex_indices",
            "property float quality", "element edge 1", "property int vertex1",
            "property int vertex2", "end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if form == "ascii":
            for p in v:
                f.write(("%s %s %s 7\n" % tuple(fmt(c) for c in p)).encode())
            for tri in t:
                f.write(("3 %d %d %d 0.5\n" % tuple(int(i) for i in tri)).encode())
            f.write(b"0 1\n")
        else:
            e = "<" if form == "binary_little_endian" else ">"
            for p in v:
                f.write(struct.pack(e + "fffB", *map(float, p), 7))
            for tri in t:
                f.write(struct.pack(e + "Biiif", 3, *map(int, tri), 0.5))
            f.write(struct.pack(e + "ii", 0, 1))


def write_stl(path, v, t, binary):
    tri = v[t]
    if binary:
        with open(path, "wb") as f:
            f.write(b"solid-looking header is fine in binary".ljust(80, b" "))
            f.write(struct.pack("<I", len(t)))
            for c in tri:
                f.write(struct.pack("<12fH", 0.0, 0.0, 0.0, *map(float, c.reshape(-1)), 0))
    else:
        with open(path, "w") as f:
            f.write("solid room\n")
            for c in tri:
                f.write("  facet normal 0 0 0\n    outer loop\n")
                for p in c:
                    f.write("      vertex %s %s %s\n" % tuple(fmt(x) for x in p))
                f.write("    endloop\n  endfacet\n")
            f.write("endsolid room\n")


def write_vtk(path, v, t, binary, v51=False):
    with open(path, "wb") as f:
        ver = "5.1" if v51 else "4.2"
        f.write(f"# vtk DataFile Version {ver}\nroom\n{'BINARY' if binary else 'ASCII'}\n"
                f"DATASET POLYDATA\nPOINTS {len(v)} float\n".encode())
        if binary:
            f.write(np.asarray(v, ">f4").tobytes())
            f.write(b"\n")
        else:
            for p in v:
                f.write(("%s %s %s\n" % tuple(fmt(c) for c in p)).encode())
        if v51:
            f.write(f"POLYGONS {len(t) + 1} {3 * len(t)}\nOFFSETS vtktypeint64\n".encode())
            off = np.arange(len(t) + 1, dtype=np.int64) * 3
            con = np.asarray(t, np.int64).reshape(-1)
            if binary:
                f.write(off.astype(">i8").tobytes() + b"\n")
                f.write(b"CONNECTIVITY vtktypeint64\n")
                f.write(con.astype(">i8").tobytes() + b"\n")
            else:
                f.write((" ".join(map(str, off)) + "\n").encode())
                f.write(b"CONNECTIVITY vtktypeint64\n")
                f.write((" ".join(map(str, con)) + "\n").encode())
        else:
            f.write(f"POLYGONS {len(t)} {4 * len(t)}\n".encode())
            cells = np.concatenate([np.full((len(t), 1), 3, np.int64), t.astype(np.int64)], 1)
            if binary:
                f.write(cells.astype(">i4").tobytes() + b"\n")
            else:
                for c in cells:
                    f.write((" ".join(map(str, c)) + "\n").encode())
        f.write(f"POINT_DATA {len(v)}\nSCALARS s float 1\nLOOKUP_TABLE default\n".encode())


def test_obj_round_trip(tmp_path, mesh):
    v, t = mesh
    p = str(tmp_path / "room.obj")
    write_obj(p, v, t)
    gv, gt = va.load_mesh(p)
    assert np.array_equal(gv, v) and np.array_equal(gt, t)


@pytest.mark.parametrize("form", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_round_trip(tmp_path, mesh, form):
    v, t = mesh
    p = str(tmp_path / "room.ply")
    write_ply(p, v, t, form)
    gv, gt = va.load_mesh(p)
    assert np.array_equal(gv, v) and np.array_equal(gt, t)


@pytest.mark.parametrize("binary,v51", [(False, False), (True, False), (False, True), (True, True)])
def test_vtk_round_trip(tmp_path, mesh, binary, v51):
    v, t = mesh
    p = str(tmp_path / "room.vtk")
    write_vtk(p, v, t, binary, v51)
    gv, gt = va.load_mesh(p)
    assert np.array_equal(gv, v) and np.array_equal(gt, t)


def first_appearance_merge(v, t):
    """Coincident corners merged, numbered by first appearance; collapsed
    facets dropped (numpy, independent of the reader)."""
    corners = v[t].reshape(-1, 3) + np.float32(0)  # -0 -> +0
    _, first, inv = np.unique(corners, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    ids = rank[inv.reshape(-1)].reshape(-1, 3)
    keep = (ids[:, 0] != ids[:, 1]) & (ids[:, 0] != ids[:, 2]) & (ids[:, 1] != ids[:, 2])
    return corners[first[order]], ids[keep].astype(np.uint32)


@pytest.mark.parametrize("binary", [False, True])
def test_stl_merges_coincident_corners(tmp_path, mesh, binary):
    v, t = mesh
    # a degenerate facet (two equal corners) and a duplicate point at -0.0
    v2 = np.concatenate([v, [[0.0, 0.0, -0.0]]]).astype(np.float32)
    t2 = np.concatenate([t, [[0, 0, 1], [len(v), 1, 2]]]).astype(np.uint32)
    p = str(tmp_path / "room.stl")
    write_stl(p, v2, t2, binary)
    gv, gt = va.load_mesh(p)
    ev, et = first_appearance_merge(v2, t2)
    assert np.array_equal(gv, ev) and np.array_equal(gt, et)
    assert gt.shape[0] == t2.shape[0] - 1  # only the degenerate facet went
    # the geometry of every kept facet is the written one
    assert np.array_equal(gv[gt], v2[t2[[i for i in range(len(t2)) if i != len(t)]]])


def test_obj_hand_written(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("# quad, pentagon, negative and slashed indices\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0 1.0\nv 0 1 0\nv 0.5 2 0\n"
                 "vt 0 0\ng part\nusemtl x\ns off\n"
                 "f 1/1 2/1 3/1 4/1\n"
                 "f -5 -4 -3 -1 -2\n"
                 "l 1 2\n")
    v, t = va.load_mesh(str(p))
    assert v.shape == (5, 3) and v[4].tolist() == [0.5, 2.0, 0.0]
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 2], [0, 2, 4], [0, 4, 3]]


def test_vtk_hand_written_skips_other_cells(tmp_path):
    p = tmp_path / "c.vtk"
    p.write_text("# vtk DataFile Version 3.0\nhand\nASCII\nDATASET POLYDATA\n"
                 "POINTS 5 double\n0 0 0 1 0 0 1 1 0\n0 1 0 0.25 0.5 1\n"
                 "VERTICES 1 2\n1 4\nLINES 1 3\n2 0 1\n"
                 "POLYGONS 2 9\n4 0 1 2 3\n3 0 1 4\n"
                 "TRIANGLE_STRIPS 1 5\n4 0 1 2 3\n"
                 "CELL_DATA 2\nSCALARS m int 1\nLOOKUP_TABLE default\n1 2\n")
    v, t = va.load_mesh(str(p))
    assert v.shape == (5, 3) and v[4].tolist() == [0.25, 0.5, 1.0]
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4]]


def test_errors(tmp_path):
    with pytest.raises(va.VoxError):
        va.load_mesh(str(tmp_path / "missing.obj"))
    p = tmp_path / "m.xyz"
    p.write_text("1 2 3\n")
    with pytest.raises(va.VoxError) as e:
        va.load_mesh(str(p))
    assert "extension" in str(e.value)
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(va.VoxError) as e:
        va.load_mesh(str(p))
    assert "out of range" in str(e.value)
    p = tmp_path / "bad.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\n"
                 "property float z\nend_header\n0 0 0\n")
    with pytest.raises(va.VoxError) as e:
        va.load_mesh(str(p))
    assert "truncated" in str(e.value)


def read_structured_points(path):
    """A tiny legacy VTK STRUCTURED_POINTS reader: dims, origin, spacing and
    the point scalars by name."""
    raw = open(path, "rb").read()
    lines = raw.split(b"\n")
    binary = lines[2].strip() == b"BINARY"
    pos = sum(len(x) + 1 for x in lines[:3])
    hdr = {}
    out = {}
    n = None
    while pos < len(raw):
        end = raw.index(b"\n", pos)
        words = raw[pos:end].split()
        pos = end + 1
        if not words:
            continue
        kw = words[0].decode()
        if kw in ("DATASET",):
            continue
        if kw in ("DIMENSIONS", "ORIGIN", "SPACING"):
            hdr[kw] = tuple(float(w) for w in words[1:])
        elif kw == "POINT_DATA":
            n = int(words[1])
        elif kw == "SCALARS":
            name, typ = words[1].decode(), words[2].decode()
            end = raw.index(b"\n", pos)  # LOOKUP_TABLE line
            pos = end + 1
            dt = {"unsigned_char": ">u1", "unsigned_short": ">u2", "float": ">f4"}[typ]
            if binary:
                size = np.dtype(dt).itemsize * n
                out[name] = np.frombuffer(raw[pos:pos + size], dt).copy()
                pos += size + 1
            else:
                vals = []
                while len(vals) < n:
                    end = raw.index(b"\n", pos)
                    vals.extend(raw[pos:end].split())
                    pos = end + 1
                out[name] = np.array([float(x) for x in vals]).astype(dt)
    return hdr, out


@pytest.mark.parametrize("kind", [va.NODE_SHORT, va.NODE_LONG, va.NODE_SHORT_FCC,
                                  va.NODE_LONG_FCC, va.NODE_PARTIAL])
@pytest.mark.parametrize("binary", [False, True])
def test_node_dump(tmp_path, kind, binary):
    rng = np.random.default_rng(kind)
    dims = (6, 4, 3)
    n = 6 * 4 * 3
    nodes = rng.integers(0, 256, n * va.NODE_SIZE[kind], dtype=np.uint8)
    if kind == va.NODE_PARTIAL:
        rec = nodes.reshape(n, 8)
        rec[:, 2:4] = 0
        rec[:, 4:] = np.frombuffer(rng.random(n, np.float32).tobytes(), np.uint8).reshape(n, 4)
    if kind == va.NODE_LONG_FCC:
        nodes.reshape(n, 4)[:, 3] = 0
    p = str(tmp_path / "n.vtk")
    va.write_nodes_vtk(p, nodes, kind, dims, (1.0, -2.0, 0.5), 0.25, binary)
    hdr, sc = read_structured_points(p)
    assert hdr["DIMENSIONS"] == dims and hdr["ORIGIN"] == (1.0, -2.0, 0.5)
    assert hdr["SPACING"] == (0.25, 0.25, 0.25)
    if kind == va.NODE_SHORT:
        bid, mat = nodes & 31, nodes >> 5
    elif kind in (va.NODE_LONG, va.NODE_PARTIAL):
        r = nodes.reshape(n, -1)
        bid, mat = r[:, 0], r[:, 1]
    elif kind == va.NODE_SHORT_FCC:
        w = nodes.view(np.uint16)
        bid, mat = w & 4095, w >> 12
    else:
        r = nodes.reshape(n, 4)
        bid, mat = r[:, :2].copy().view(np.uint16).reshape(-1), r[:, 2]
    assert np.array_equal(sc["bid"].astype(np.int64), bid.astype(np.int64))
    assert np.array_equal(sc["material"].astype(np.int64), mat.astype(np.int64))
    if kind == va.NODE_PARTIAL:
        ratio = nodes.reshape(n, 8)[:, 4:].copy().view(np.float32).reshape(-1)
        assert np.array_equal(sc["ratio"].astype(np.float32), ratio)


@pytest.mark.parametrize("binary", [False, True])
def test_bits_dump(tmp_path, binary):
    rng = np.random.default_rng(5)
    dims = (64, 3, 2)
    bits = rng.integers(0, 2 ** 32, 2 * 3 * 2, dtype=np.uint64).astype(np.uint32)
    p = str(tmp_path / "b.vtk")
    va.write_bits_vtk(p, bits, dims, binary=binary)
    _, sc = read_structured_points(p)
    # voxel x of a word at bit 31 - x % 32
    exp = np.unpackbits(bits.astype(">u4").view(np.uint8))
    assert np.array_equal(sc["solid"].astype(np.uint8), exp)


@pytest.mark.gpu
def test_loaded_mesh_voxelizes_like_the_oracle(tmp_path, gpu_ctx, oracle, mesh):
    v, t = mesh
    p = str(tmp_path / "room.ply")
    write_ply(p, v, t, "binary_little_endian")
    gv, gt = va.load_mesh(p)
    gpu_ctx.set_mesh(gv, gt)
    g = gpu_ctx.plan_grid(va.PLAN_CUBE_LENGTH, 0.02)
    got = gpu_ctx.voxelize_nodes(g, va.NODE_LONG, False)
    m = W.Mesh(v, t, None, "room")
    ref = oracle.voxelize_nodes(m, oracle.plan_grid(v, va.PLAN_CUBE_LENGTH, 0.02), va.NODE_LONG,
                                False)
    assert np.array_equal(got, ref)
