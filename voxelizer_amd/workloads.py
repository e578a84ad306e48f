"""Seeded procedural meshes for the voxelizer workloads.

All meshes are closed, indexed triangle meshes with FP32 vertices and uint32
indices (the reference's only model format, include/voxelizer.h:115-130).
Randomness comes from numpy's PCG64 with an explicit seed; transforms use only
+, -, *, / and sqrt (IEEE correctly rounded), Gaussian samples are Irwin-Hall
sums of 12 uniforms added in a fixed order, so a seed gives the same buffers on
any host with this numpy.

Workloads (sizes from the reference's use case: acoustic FDTD rooms):
  icosphere()       1280 triangles, r=0.4 in the unit cube, N(0,1e-3) jitter
  shoebox_room()    5x4x3 m box + 64 rotated interior boxes, 19212 triangles
  hall()            ~1M triangles, bbox exactly (1533,1021,765)/64 m
  batch_meshes()    many small meshes of uneven size
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float32
    indices: np.ndarray   # (T, 3) uint32
    materials: np.ndarray | None = None  # (T,) uint8
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def nt(self) -> int:
        return int(self.indices.shape[0])

    def digest(self) -> str:
        import hashlib
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.vertices, np.float32).tobytes())
        h.update(np.ascontiguousarray(self.indices, np.uint32).tobytes())
        if self.materials is not None:
            h.update(np.ascontiguousarray(self.materials, np.uint8).tobytes())
        return h.hexdigest()[:16]


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def _gauss(rng: np.random.Generator, n: int) -> np.ndarray:
    """Irwin-Hall(12) - 6: unit-variance, added column by column."""
    u = rng.random((n, 12))
    s = u[:, 0].copy()
    for k in range(1, 12):
        s = s + u[:, k]
    return s - 6.0


def _unit_quat(rng: np.random.Generator) -> np.ndarray:
    while True:
        q = _gauss(rng, 4)
        n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]
        if n2 > 1e-6:
            return q / math.sqrt(n2)


def _quat_to_mat(q: np.ndarray) -> np.ndarray:
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ], dtype=np.float64)


class _Builder:
    """Concatenates closed parts into one indexed mesh."""

    def __init__(self):
        self.v = []
        self.t = []
        self.m = []
        self.nv = 0

    def add(self, verts: np.ndarray, tris: np.ndarray, mats=None):
        self.v.append(np.asarray(verts, np.float64))
        self.t.append(np.asarray(tris, np.int64) + self.nv)
        if mats is not None:
            self.m.append(np.asarray(mats, np.uint8))
        self.nv += verts.shape[0]

    def mesh(self, name: str, **meta) -> Mesh:
        v = np.concatenate(self.v).astype(np.float32)
        t = np.concatenate(self.t).astype(np.uint32)
        m = np.concatenate(self.m) if self.m else None
        return Mesh(v, t, m, name, dict(meta))


# ---------------------------------------------------------------------------
# Primitive closed surfaces
# ---------------------------------------------------------------------------

_ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)


def _normalize_rows(v: np.ndarray) -> np.ndarray:
    n = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    return v / n[:, None]


def unit_icosphere(subdiv: int):
    """Icosahedron subdivided `subdiv` times, vertices on the unit sphere."""
    p = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
                  [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    verts = list(_normalize_rows(v))
    faces = _ICO_FACES.copy()
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = (verts[a] + verts[b]) * 0.5
                m = m / math.sqrt(float(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]))
                cache[key] = len(verts)
                verts.append(m)
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = np.array(nf, dtype=np.int64)
    return np.array(verts), faces


_BOX_CACHE: dict = {}


def unit_box_surface(nx: int, ny: int, nz: int):
    """Surface lattice of [-0.5,0.5]^3 with nx*ny*nz quads per axis pair,
    shared vertices along edges, outward triangles."""
    key = (nx, ny, nz)
    if key in _BOX_CACHE:
        return _BOX_CACHE[key]
    idx = {}
    verts = []

    def vid(i, j, k):
        kk = (i, j, k)
        if kk not in idx:
            idx[kk] = len(verts)
            verts.append((i / nx - 0.5, j / ny - 0.5, k / nz - 0.5))
        return idx[kk]

    tris = []

    def face(n1, n2, fix_axis, fix_val, flip):
        for a in range(n1):
            for b in range(n2):
                q = []
                for da, db in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    c = [0, 0, 0]
                    free = [ax for ax in range(3) if ax != fix_axis]
                    c[free[0]] = a + da
                    c[free[1]] = b + db
                    c[fix_axis] = fix_val
                    q.append(vid(*c))
                if flip:
                    q = q[::-1]
                tris.append((q[0], q[1], q[2]))
                tris.append((q[0], q[2], q[3]))

    face(ny, nz, 0, 0, True)
    face(ny, nz, 0, nx, False)
    face(nx, nz, 1, 0, False)
    face(nx, nz, 1, ny, True)
    face(nx, ny, 2, 0, True)
    face(nx, ny, 2, nz, False)
    out = (np.array(verts, np.float64), np.array(tris, np.int64))
    _BOX_CACHE[key] = out
    return out


def prism_surface(n: int, nz: int):
    """Closed n-gon prism, radius 1, z in [0,1], nz vertical segments."""
    ang = np.arange(n) * (2.0 * math.pi / n)
    # cos/sin of a fixed grid of angles: computed once per n from the
    # rotation recurrence (only + and *) so hosts agree bit for bit.
    c0, s0 = math.cos(2.0 * math.pi / n), math.sin(2.0 * math.pi / n)
    cs = np.empty((n, 2))
    cs[0] = (1.0, 0.0)
    for i in range(1, n):
        cs[i, 0] = cs[i - 1, 0] * c0 - cs[i - 1, 1] * s0
        cs[i, 1] = cs[i - 1, 0] * s0 + cs[i - 1, 1] * c0
    del ang
    verts = []
    for k in range(nz + 1):
        for i in range(n):
            verts.append((cs[i, 0], cs[i, 1], k / nz))
    bot_c = len(verts)
    verts.append((0.0, 0.0, 0.0))
    top_c = len(verts)
    verts.append((0.0, 0.0, 1.0))
    tris = []
    for k in range(nz):
        for i in range(n):
            a = k * n + i
            b = k * n + (i + 1) % n
            c = (k + 1) * n + (i + 1) % n
            d = (k + 1) * n + i
            tris += [(a, b, c), (a, c, d)]
    for i in range(n):
        tris.append((bot_c, (i + 1) % n, i))
        tris.append((top_c, nz * n + i, nz * n + (i + 1) % n))
    return np.array(verts, np.float64), np.array(tris, np.int64)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def icosphere(seed: int = 0, radius: float = 0.4, subdiv: int = 3,
              jitter: float = 1e-3) -> Mesh:
    """Closed icosphere (1280 triangles at subdiv 3) centred in the unit cube."""
    rng = _rng(seed)
    v, f = unit_icosphere(subdiv)
    v = v * radius + 0.5
    v = v + jitter * _gauss(rng, v.shape[0] * 3).reshape(-1, 3)
    b = _Builder()
    b.add(v, f)
    m = b.mesh("icosphere", seed=seed, radius=radius, subdiv=subdiv)
    m.materials = np.ones(m.nt, np.uint8)
    return m


def room_box(lx: float, ly: float, lz: float):
    v = np.array([[0, 0, 0], [lx, 0, 0], [lx, ly, 0], [0, ly, 0],
                  [0, 0, lz], [lx, 0, lz], [lx, ly, lz], [0, ly, lz]], np.float64)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7],
                  [0, 1, 5], [0, 5, 4], [1, 2, 6], [1, 6, 5],
                  [2, 3, 7], [2, 7, 6], [3, 0, 4], [3, 4, 7]], np.int64)
    return v, t


def shoebox_room(seed: int = 1, n_boxes: int = 64, size=(5.0, 4.0, 3.0),
                 subdiv: int = 5, n_materials: int = 8,
                 reject_overlaps: bool = True) -> Mesh:
    """5x4x3 m room + n_boxes rotated boxes (sides U[0.2,1.5] m), each face
    subdivided subdiv x subdiv; materials U{1..n_materials} per triangle."""
    rng = _rng(seed)
    b = _Builder()
    rv, rt = room_box(*size)
    b.add(rv, rt)
    uv, ut = unit_box_surface(subdiv, subdiv, subdiv)
    placed = []
    for _ in range(n_boxes):
        for _attempt in range(200):
            s = 0.2 + 1.3 * rng.random(3)
            r = 0.5 * math.sqrt(float(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]))
            c = np.array([r + (size[a] - 2 * r) * rng.random() for a in range(3)])
            R = _quat_to_mat(_unit_quat(rng))
            ok = True
            if reject_overlaps:
                for (c2, r2) in placed:
                    dd = c - c2
                    if dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < (r + r2) ** 2:
                        ok = False
                        break
            if ok:
                break
        placed.append((c, r))
        v = (uv * s[None, :]) @ R.T + c[None, :]
        b.add(v, ut)
    m = b.mesh("shoebox_room", seed=seed, n_boxes=n_boxes)
    m.materials = (1 + rng.integers(0, n_materials, m.nt)).astype(np.uint8)
    return m


def _tri_wave(t):
    return 4.0 * np.abs(t - np.floor(t + 0.5)) - 1.0


def hall(seed: int = 2, n_seats: int = 1600, n_columns: int = 83,
         n_ornaments: int = 317, n_materials: int = 8, scale: float = 1.0) -> Mesh:
    """Large hall: heightfield floor, barrel-vaulted ceiling, walls, seats
    (subdivided boxes), columns (prisms) and hanging ornaments (icospheres).
    Subdivision levels are log-normal, so triangle areas are heavy tailed.
    bbox = [0,1533/64] x [0,1021/64] x [0,765/64] m exactly (scale=1).
    Materials are 0..n_materials-1 (ShortNode holds 3 material bits)."""
    rng = _rng(seed)
    Lx, Ly, Lz = 1533.0 / 64.0 * scale, 1021.0 / 64.0 * scale, 765.0 / 64.0 * scale
    H_side = 8.0 * scale
    nx, ny = 96, 64
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = np.linspace(0.0, Ly, ny + 1)
    xs[-1], ys[-1] = Lx, Ly
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    zf = 0.15 * scale * (1.0 + _tri_wave(X / (3.0 * scale)) * _tri_wave(Y / (2.0 * scale)))
    zf = zf - zf.min()
    cy = ys[ny // 2]
    ceil_ = H_side + (Lz - H_side) * (1.0 - ((Y - cy) / (Ly / 2.0)) ** 2)
    ceil_[:, ny // 2] = Lz
    nv_w = 24
    # Vertex layout: floor grid, ceiling grid, wall interior rings.
    fid = lambda i, j: i * (ny + 1) + j
    nf = (nx + 1) * (ny + 1)
    verts = [np.stack([X.ravel(), Y.ravel(), zf.ravel()], 1),
             np.stack([X.ravel(), Y.ravel(), ceil_.ravel()], 1)]
    cid = lambda i, j: nf + fid(i, j)
    tris = []
    for i in range(nx):
        for j in range(ny):
            a, bb, c, d = fid(i, j), fid(i + 1, j), fid(i + 1, j + 1), fid(i, j + 1)
            tris += [(a, c, bb), (a, d, c)]
            a, bb, c, d = cid(i, j), cid(i + 1, j), cid(i + 1, j + 1), cid(i, j + 1)
            tris += [(a, bb, c), (a, c, d)]
    perim = ([(i, 0) for i in range(nx)] + [(nx, j) for j in range(ny)] +
             [(i, ny) for i in range(nx, 0, -1)] + [(0, j) for j in range(ny, 0, -1)])
    base = 2 * nf
    ring = []
    for p, (i, j) in enumerate(perim):
        z0, z1 = zf[i, j], ceil_[i, j]
        col = [fid(i, j)]
        for k in range(1, nv_w):
            verts.append(np.array([[xs[i], ys[j], z0 + (z1 - z0) * (k / nv_w)]]))
            col.append(base)
            base += 1
        col.append(cid(i, j))
        ring.append(col)
    for p in range(len(perim)):
        c0, c1 = ring[p], ring[(p + 1) % len(perim)]
        for k in range(nv_w):
            tris += [(c0[k], c1[k], c1[k + 1]), (c0[k], c1[k + 1], c0[k + 1])]
    bld = _Builder()
    bld.add(np.concatenate(verts), np.array(tris, np.int64))
    # Seat lattice (pitch 0.45 m, footprint <= 0.3 m: any yaw is overlap free).
    pitch = 0.45 * scale
    gx = np.arange(0.6 * scale, Lx - 0.6 * scale, pitch)
    gy = np.arange(0.6 * scale, Ly - 0.6 * scale, pitch)
    cells = [(a, b) for a in gx for b in gy]
    order = rng.permutation(len(cells))
    floor_top = float(zf.max()) + 0.02 * scale

    def lognormal_level(mu, sigma, lo, hi):
        g = float(_gauss(rng, 1)[0])
        return int(min(hi, max(lo, round(math.exp(mu + sigma * g)))))

    k = 0
    for _ in range(n_seats):
        cx, cyy = cells[order[k % len(cells)]]
        k += 1
        w = (0.22 + 0.08 * rng.random()) * scale
        dpt = (0.22 + 0.08 * rng.random()) * scale
        hgt = (0.4 + 0.5 * rng.random()) * scale
        lvl = lognormal_level(math.log(2.8), 0.75, 1, 40)
        uv, ut = unit_box_surface(lvl, lvl, max(1, lvl * 2))
        ang = _gauss(rng, 2)
        nrm = math.sqrt(float(ang[0] * ang[0] + ang[1] * ang[1])) or 1.0
        ca, sa = ang[0] / nrm, ang[1] / nrm
        R = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
        v = (uv * np.array([w, dpt, hgt])) @ R.T
        v = v + np.array([cx, cyy, floor_top + hgt / 2.0])
        bld.add(v, ut)
    for _ in range(n_columns):
        cx, cyy = cells[order[k % len(cells)]]
        k += 1
        nseg = lognormal_level(math.log(16.0), 0.75, 6, 64)
        nz = lognormal_level(math.log(20.0), 0.75, 4, 200)
        pv, pt = prism_surface(nseg, nz)
        rad = 0.13 * scale
        hgt = 7.5 * scale - floor_top
        v = pv * np.array([rad, rad, hgt]) + np.array([cx, cyy, floor_top])
        bld.add(v, pt)
    for _ in range(n_ornaments):
        cx, cyy = cells[order[k % len(cells)]]
        k += 1
        lvl = lognormal_level(math.log(2.0), 0.75, 0, 4)
        iv, it = unit_icosphere(lvl)
        rad = (0.08 + 0.04 * rng.random()) * scale
        v = iv * rad + np.array([cx, cyy, 7.75 * scale])
        bld.add(v, it)
    m = bld.mesh("hall", seed=seed)
    m.materials = rng.integers(0, n_materials, m.nt).astype(np.uint8)
    return m


def batch_meshes(n: int = 256, seed: int = 4, tri_lo: int = 500,
                 tri_hi: int = 200_000):
    """n meshes with triangle counts roughly log-uniform in [tri_lo, tri_hi]:
    icospheres, rooms with boxes and subdivided boxes, each with a cube length
    that gives 64..512 voxels along its longest side."""
    rng = _rng(seed)
    out = []
    for i in range(n):
        target = math.exp(math.log(tri_lo) + (math.log(tri_hi) - math.log(tri_lo)) * rng.random())
        kind = i % 3
        if kind == 0:
            lvl = max(2, min(7, int(round(math.log(target / 20.0) / math.log(4.0)))))
            m = icosphere(seed=seed * 1000 + i, subdiv=lvl)
        elif kind == 1:
            nb = max(1, min(64, int(target // 300)))
            sub = max(1, min(20, int(round(math.sqrt(max(1.0, target / (12.0 * nb)))))))
            m = shoebox_room(seed=seed * 1000 + i, n_boxes=nb, subdiv=sub)
        else:
            s = max(1, int(round(math.sqrt(target / 12.0))))
            uv, ut = unit_box_surface(s, s, s)
            sc = 0.5 + rng.random(3)
            b = _Builder()
            b.add(uv * sc + 1.0, ut)
            m = b.mesh("box", seed=seed * 1000 + i)
            m.materials = (1 + rng.integers(0, 8, m.nt)).astype(np.uint8)
        ext = (m.vertices.max(0) - m.vertices.min(0)).astype(np.float64)
        longest = 64 + int(rng.integers(0, 449))
        m.meta["cube_length"] = float(ext.max()) / float(longest)
        out.append(m)
    return out


def tetrahedron(scale=1.0, offset=(0.0, 0.0, 0.0)) -> Mesh:
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64) * scale + offset
    t = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.int64)
    b = _Builder()
    b.add(v, t)
    m = b.mesh("tetrahedron")
    m.materials = np.array([1, 2, 3, 4], np.uint8)
    return m


def box(lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), subdiv=1) -> Mesh:
    uv, ut = unit_box_surface(subdiv, subdiv, subdiv)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    v = (uv + 0.5) * (hi - lo) + lo
    b = _Builder()
    b.add(v, ut)
    m = b.mesh("box")
    m.materials = (1 + (np.arange(m.nt) % 7)).astype(np.uint8)
    return m


def random_soup(seed: int, n_tris: int, lo=0.0, hi=1.0) -> Mesh:
    """Open triangle soup (not closed): exercises odd-parity rows."""
    rng = _rng(seed)
    v = lo + (hi - lo) * rng.random((3 * n_tris, 3))
    t = np.arange(3 * n_tris, dtype=np.int64).reshape(-1, 3)
    b = _Builder()
    b.add(v, t)
    m = b.mesh("soup", seed=seed)
    m.materials = (1 + rng.integers(0, 8, m.nt)).astype(np.uint8)
    return m


def normal_x_signs(v: np.ndarray):
    """(tN, n.x) of triangles v (T, 3, 3) in float64, in the two expressions
    the reference uses: the solid filter's tN (src/voxelizer.cu:882-884) and
    the x component of cross(v0 - v2, v1 - v0) of the fill
    (src/voxelizer.cu:1375-1378).  Both equal the same exact value; rounding
    can give them different signs."""
    v = np.asarray(v, np.float64)
    y, z = v[:, :, 1], v[:, :, 2]
    tN = (y[:, 0] * z[:, 1] + y[:, 2] * z[:, 0] + z[:, 2] * y[:, 1]) - \
         (y[:, 2] * z[:, 1] + z[:, 0] * y[:, 1] + z[:, 2] * y[:, 0])
    nx = (y[:, 0] - y[:, 2]) * (z[:, 1] - z[:, 0]) - (z[:, 0] - z[:, 2]) * (y[:, 1] - y[:, 0])
    return tN, nx


def disagreeing_slivers(seed: int = 3, n: int = 64, box_hi=(8.0, 34.0, 34.0)) -> Mesh:
    """A closed box [0, box_hi] plus n triangles whose y/z projection is a
    segment through the origin (vertices (x, a s, b s) with small odd a, b and
    s of at most 20 significant bits, so every coordinate is exact in FP32)
    and whose tN and n.x have different signs: the sums in tN round where the
    exact value is 0, n.x is computed exactly.  Such triangles orient the
    solid filter's edges one way and the fill's the other."""
    rng = _rng(seed)
    dirs = [(3, 5), (5, 3), (3, 7), (7, 3), (1, 3), (3, 1)]
    tris = []
    while len(tris) < n:
        a, b = dirs[int(rng.integers(0, len(dirs)))]
        lim = min(box_hi[1] / a, box_hi[2] / b) * 0.95

        def few_bits(lo, hi):
            x = rng.uniform(lo, hi)
            e = math.floor(math.log2(x))
            return round(x / 2.0 ** (e - 19)) * 2.0 ** (e - 19)
        s = [few_bits(2.0 ** -30, 2.0 ** -29), few_bits(0.55 * lim, lim),
             few_bits(0.05 * lim, 0.5 * lim)]
        if rng.random() < 0.5:
            s[1], s[2] = s[2], s[1]
        xs = rng.uniform(0.5, box_hi[0] - 0.5, 3).astype(np.float32).astype(np.float64)
        t = np.array([[xs[k], a * s[k], b * s[k]] for k in range(3)], np.float64)
        assert np.array_equal(t.astype(np.float32).astype(np.float64), t)
        tN, nx = normal_x_signs(t[None])
        if tN[0] != 0.0 and (tN[0] > 0.0) != (nx[0] > 0.0):
            tris.append(t)
    bx = box((0.0, 0.0, 0.0), box_hi, subdiv=2)
    b = _Builder()
    b.add(bx.vertices.astype(np.float64), bx.indices.astype(np.int64), bx.materials)
    v = np.concatenate(tris)
    b.add(v, np.arange(v.shape[0], dtype=np.int64).reshape(-1, 3),
          (1 + np.arange(n) % 7).astype(np.uint8))
    return b.mesh("slivers", seed=seed)
