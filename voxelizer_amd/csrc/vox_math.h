// vox_math.h -- device-side arithmetic of the voxelizer, in the operation
// order of the reference's expressions (FP64 solid stage, FP32 surface stage).
// Compiled with -ffp-contract=off: no FMA contraction anywhere.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VOX_DBL_MIN 2.2250738585072014e-308

namespace vk {

struct D2 { double x, y; };
struct D3 { double x, y, z; };
struct F2 { float x, y; };
struct F3 { float x, y, z; };
struct I3 { int32_t x, y, z; };

__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 fsub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// cross(a, b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ F3 fcross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// dot evaluated left to right
__device__ __forceinline__ float fdot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Saturating conversions (truncate toward zero, clamp, NaN -> 0), the
// semantics of the hardware float->int conversions on both vendors.
__device__ __forceinline__ int32_t sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483648.0) return INT32_MAX;
    if (v <= -2147483649.0) return INT32_MIN;
    return (int32_t)v;
}
__device__ __forceinline__ uint32_t sat_u32(double v) {
    if (v != v) return 0u;
    if (v >= 4294967296.0) return UINT32_MAX;
    if (v <= -1.0) return 0u;
    return (uint32_t)v;
}
// sat_i32((double)v) without the double: NaN -> 0, v >= 2^31 -> INT32_MAX,
// v < -2^31 -> INT32_MIN (the largest float below -2^31 is -2^31 - 128, so
// the double bound -2^31 - 1 selects the same floats), else truncation.
// Written as selects so that it compiles to compares, not branches.
__device__ __forceinline__ int32_t f2i(float v) {
    const int32_t t = (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : 0;
    return v != v ? 0 : v >= 2147483648.0f ? INT32_MAX : v < -2147483648.0f ? INT32_MIN : t;
}

__device__ __forceinline__ D3 fetch_d(const float *__restrict__ V, const uint32_t *__restrict__ I,
                                      uint32_t t, int k) {
    uint32_t i = I[3u * t + (uint32_t)k];
    return {(double)V[3u * i], (double)V[3u * i + 1u], (double)V[3u * i + 2u]};
}
__device__ __forceinline__ F3 fetch_f(const float *__restrict__ V, const uint32_t *__restrict__ I,
                                      uint32_t t, int k) {
    uint32_t i = I[3u * t + (uint32_t)k];
    return {V[3u * i], V[3u * i + 1u], V[3u * i + 2u]};
}

// ---------------------------------------------------------------------------
// Solid stage (FP64)
// ---------------------------------------------------------------------------

// Work-queue filter of one triangle for a y/z subspace [smin, smax):
// ext voxel range of its yz bounding box intersected with the subspace
// ("commonSpace"), its 4x4 tile range, and edge functions oriented by the
// reference's tN expression (src/voxelizer.cu:861-945).
struct TileSetup {
    D2 ne[3];
    double de[3];
    uint32_t cmin[2], cmax[2];
    uint32_t tmin[2], tmax[2];
};

// x component of the normal in the reference's own expression
// (src/voxelizer.cu:882-884) and the edge functions it orients.
__device__ __forceinline__ double tri_tN(const D3 v[3]) {
    return (v[0].y * v[1].z + v[2].y * v[0].z + v[2].z * v[1].y) -
           (v[2].y * v[1].z + v[0].z * v[1].y + v[2].z * v[0].y);
}
__device__ __forceinline__ void edges_by_sign(const D3 v[3], bool pos, TileSetup &ts) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const D3 a = v[i], b = v[(i + 1) % 3];
        if (pos)
            ts.ne[i] = {a.z - b.z, b.y - a.y};
        else
            ts.ne[i] = {b.z - a.z, a.y - b.y};
        ts.de[i] = -(ts.ne[i].x * a.y + ts.ne[i].y * a.z);
    }
}
// Edge functions only (the commonSpace box comes from the setup pass).
__device__ __forceinline__ void tile_edges(const D3 v[3], TileSetup &ts) {
    edges_by_sign(v, tri_tN(v) > 0.0, ts);
}

__device__ __forceinline__ bool tile_setup(const D3 v[3], double emy, double emz, double d,
                                           uint32_t sminy, uint32_t sminz, uint32_t smaxy,
                                           uint32_t smaxz, TileSetup &ts) {
    D2 bmin = {v[0].y, v[0].z}, bmax = bmin;
#pragma unroll
    for (int i = 1; i < 3; i++) {
        bmin.x = v[i].y < bmin.x ? v[i].y : bmin.x;
        bmin.y = v[i].z < bmin.y ? v[i].z : bmin.y;
        bmax.x = v[i].y > bmax.x ? v[i].y : bmax.x;
        bmax.y = v[i].z > bmax.y ? v[i].z : bmax.y;
    }
    double vb0 = (bmin.x - emy) / d, vb1 = (bmin.y - emz) / d;
    double vb2 = (bmax.x - emy) / d, vb3 = (bmax.y - emz) / d;
    int32_t vmin0 = sat_i32(ceil(vb0)), vmin1 = sat_i32(ceil(vb1));
    int32_t vmax0 = sat_i32(floor(vb2) + 1.0), vmax1 = sat_i32(floor(vb3) + 1.0);
    const double tN = tri_tN(v);
    if (tN == 0.0) return false;
    if (vmin0 >= vmax0 || vmin1 >= vmax1) return false;
    if (vmax0 <= (int32_t)sminy || vmax1 <= (int32_t)sminz) return false;
    if (vmin0 >= (int32_t)smaxy || vmin1 >= (int32_t)smaxz) return false;
    ts.cmin[0] = (uint32_t)max((int32_t)sminy, vmin0);
    ts.cmin[1] = (uint32_t)max((int32_t)sminz, vmin1);
    ts.cmax[0] = min(smaxy, (uint32_t)vmax0);
    ts.cmax[1] = min(smaxz, (uint32_t)vmax1);
    edges_by_sign(v, tN > 0.0, ts);
    ts.tmin[0] = ts.cmin[0] / 4u;
    ts.tmin[1] = ts.cmin[1] / 4u;
    ts.tmax[0] = (ts.cmax[0] - 1u) / 4u;
    ts.tmax[1] = (ts.cmax[1] - 1u) / 4u;
    return true;
}

// Tile (ty, tz) is queued if a voxel centre of it inside commonSpace passes
// all three (>= 0) edge tests (src/voxelizer.cu:998-1031).
__device__ __forceinline__ bool tile_hit(const TileSetup &ts, uint32_t ty, uint32_t tz,
                                         double emy, double emz, double d) {
    bool hit = false;
    for (uint32_t i = ty * 4u; i < ty * 4u + 4u && i < ts.cmax[0] && !hit; i++) {
        for (uint32_t j = tz * 4u; j < tz * 4u + 4u && j < ts.cmax[1] && !hit; j++) {
            if (i >= ts.cmin[0] && j >= ts.cmin[1]) {
                double px = double(i) * d + emy;
                double py = double(j) * d + emz;
                bool r = true;
#pragma unroll
                for (int k = 0; k < 3; k++)
                    r &= (ts.ne[k].x * px + ts.ne[k].y * py + ts.de[k] >= 0.0);
                hit = r;
            }
        }
    }
    return hit;
}

// Per-triangle data of the centre test (src/voxelizer.cu:1334-1372).
struct FillSetup {
    D3 v0, n;
    D2 ne[3];
    double de[3];
};
__device__ __forceinline__ void fill_setup(const D3 v[3], FillSetup &fs) {
    fs.v0 = v[0];
    fs.n = dcross(dsub(v[0], v[2]), dsub(v[1], v[0]));
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const D3 a = v[i], b = v[(i + 1) % 3];
        if (fs.n.x > 0.0)
            fs.ne[i] = {a.z - b.z, b.y - a.y};
        else
            fs.ne[i] = {b.z - a.z, a.y - b.y};
        fs.de[i] = -(fs.ne[i].x * a.y + fs.ne[i].y * a.z);
    }
}
// Centre test with the DBL_MIN top-left tie break and the x crossing
// (src/voxelizer.cu:1418-1453); voxelX includes the +1 x padding.
__device__ __forceinline__ bool fill_center(const FillSetup &fs, double emx, double emy,
                                            double emz, double d, uint32_t cy, uint32_t cz,
                                            uint32_t &voxelX) {
    double px = double(cy) * d + emy;
    double py = double(cz) * d + emz;
    uint32_t ok = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double r = fs.ne[i].x * px + fs.ne[i].y * py + fs.de[i];
        if (fs.ne[i].x > 0.0 || (fs.ne[i].x == 0.0 && fs.ne[i].y < 0.0)) r += VOX_DBL_MIN;
        if (r > 0.0) ok++;
    }
    if (ok != 3u) return false;
    double ax = fs.v0.x - emx, ay = fs.v0.y - px, az = fs.v0.z - py;
    double B = ax * fs.n.x + ay * fs.n.y + az * fs.n.z;
    double x = B / fs.n.x;
    x = x < 0.0 ? 0.0 : x;
    voxelX = sat_u32(ceil(x / d)) + 1u;
    return true;
}

// ---------------------------------------------------------------------------
// Surface stage (FP32)
// ---------------------------------------------------------------------------
enum { BB_DEGEN = 0, BB_1D = 1, BB_2D = 2, BB_3D = 3 };
enum { AX_X = 0, AX_Y = 1, AX_Z = 2 };

struct VoxBB { I3 min, max; };
struct OverlapData { F2 ne[3]; float de[3]; };

__device__ __forceinline__ void tri_bounds(const F3 t[3], F3 &mn, F3 &mx) {
    mn = t[0];
    mx = t[0];
#pragma unroll
    for (int i = 1; i < 3; i++) {
        if (t[i].x < mn.x) mn.x = t[i].x;
        if (t[i].y < mn.y) mn.y = t[i].y;
        if (t[i].z < mn.z) mn.z = t[i].z;
        if (t[i].x > mx.x) mx.x = t[i].x;
        if (t[i].y > mx.y) mx.y = t[i].y;
        if (t[i].z > mx.z) mx.z = t[i].z;
    }
}
// getVoxelBoundsHalf (src/voxelizer.cu:2471-2492).
__device__ __forceinline__ VoxBB voxel_bounds_half(F3 mn, F3 mx, F3 m, float d) {
    VoxBB b;
    b.min = {f2i(ceilf((mn.x - m.x) / d - 0.5f)), f2i(ceilf((mn.y - m.y) / d - 0.5f)),
             f2i(ceilf((mn.z - m.z) / d - 0.5f))};
    b.max = {f2i(floorf((mx.x - m.x) / d + 0.5f)), f2i(floorf((mx.y - m.y) / d + 0.5f)),
             f2i(floorf((mx.z - m.z) / d + 0.5f))};
    return b;
}
__device__ __forceinline__ float vox_corner(int32_t c, float m, float d) {
    return m + (float(c) - 0.5f) * d;
}

// classifyTriangles + analyzeBoundingBox: class (0..3) and dominant axis,
// from the triangle's unclipped voxel range vb (voxel_bounds_half of its
// bounds).
__device__ __forceinline__ void classify_vb(const F3 t[3], const VoxBB &vb, int &bb, int &axis) {
    F3 n = fcross(fsub(t[0], t[2]), fsub(t[1], t[0]));
    if (!(n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)) {
        bb = BB_DEGEN;
        axis = AX_X;
        return;
    }
    int32_t dx = vb.max.x - vb.min.x, dy = vb.max.y - vb.min.y, dz = vb.max.z - vb.min.z;
    if (dx == 0) {
        if (dy == 0) {
            bb = BB_1D;
            axis = dz == 0 ? AX_X : AX_Z;
        } else if (dz == 0) {
            bb = BB_1D;
            axis = AX_Y;
        } else {
            bb = BB_2D;
            axis = AX_X;
        }
    } else if (dy == 0) {
        if (dz == 0) {
            bb = BB_1D;
            axis = AX_X;
        } else {
            bb = BB_2D;
            axis = AX_Y;
        }
    } else if (dz == 0) {
        bb = BB_2D;
        axis = AX_Z;
    } else {
        bb = BB_3D;
        float ax = fabsf(n.x), ay = fabsf(n.y), az = fabsf(n.z);
        if (ax > ay && ax > az)
            axis = AX_X;
        else if (ay > ax && ay > az)
            axis = AX_Y;
        else
            axis = AX_Z;
    }
}

// setup{YZ,ZX,XY}OverlapTest (src/voxelizer.cu:2677-2757): plane 0 = YZ
// (coordinates (y,z)), 1 = ZX ((z,x)), 2 = XY ((x,y)).
__device__ __forceinline__ void setup_overlap(const F3 t[3], F3 n, float d, int plane,
                                              OverlapData &o) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const F3 a = t[i], b = t[(i + 1) % 3];
        float a0, a1, b0, b1;  // projected (first, second) coordinates
        bool pos;
        if (plane == 0) {
            a0 = a.y; a1 = a.z; b0 = b.y; b1 = b.z; pos = n.x >= 0.0f;
        } else if (plane == 1) {
            a0 = a.z; a1 = a.x; b0 = b.z; b1 = b.x; pos = n.y >= 0.0f;
        } else {
            a0 = a.x; a1 = a.y; b0 = b.x; b1 = b.y; pos = n.z >= 0.0f;
        }
        if (pos)
            o.ne[i] = {a1 - b1, b0 - a0};
        else
            o.ne[i] = {b1 - a1, a0 - b0};
        o.de[i] = -(o.ne[i].x * a0 + o.ne[i].y * a1) + fmaxf(0.0f, d * o.ne[i].x) +
                  fmaxf(0.0f, d * o.ne[i].y);
    }
}
__device__ __forceinline__ bool overlap2(const OverlapData &o, float p0, float p1) {
    bool r = true;
#pragma unroll
    for (int i = 0; i < 3; i++) r &= (o.ne[i].x * p0 + o.ne[i].y * p1) + o.de[i] >= 0.0f;
    return r;
}

// determineDepthRange{X,Y,Z} (src/voxelizer.cu:4004-4119) for a voxel column
// whose corner is p along the dominant axis `axis`.
__device__ __forceinline__ void depth_range(const F3 t0, F3 p, F3 n, F3 m, float d, int axis,
                                            int32_t &lo, int32_t &hi) {
    F3 pmin, pmax;
    if (axis == AX_X) {
        bool ny = n.y > 0.0f, nz = n.z > 0.0f;
        pmin = {0.0f, p.y + float(ny) * d, p.z + float(nz) * d};
        pmax = {0.0f, p.y + float(!ny) * d, p.z + float(!nz) * d};
    } else if (axis == AX_Y) {
        bool nx = n.x > 0.0f, nz = n.z > 0.0f;
        pmin = {p.x + float(nx) * d, 0.0f, p.z + float(nz) * d};
        pmax = {p.x + float(!nx) * d, 0.0f, p.z + float(!nz) * d};
    } else {
        bool nx = n.x > 0.0f, ny = n.y > 0.0f;
        pmin = {p.x + float(nx) * d, p.y + float(ny) * d, 0.0f};
        pmax = {p.x + float(!nx) * d, p.y + float(!ny) * d, 0.0f};
    }
    float na = axis == AX_X ? n.x : axis == AX_Y ? n.y : n.z;
    float a = fdot3(fsub(t0, pmin), n) / na;
    float b = fdot3(fsub(t0, pmax), n) / na;
    if (a > b) {
        float tmp = a;
        a = b;
        b = tmp;
    }
    float mc = axis == AX_X ? m.x : axis == AX_Y ? m.y : m.z;
    lo = f2i(ceilf((a - mc) / d - 0.5f));
    hi = f2i(floorf((b - mc) / d + 0.5f));
}

// ---------------------------------------------------------------------------
// PartialNode volume ratio: processVoxel<PartialNode> (src/voxelizer.cu:
// 2997-3033) with calculateVoxelPlaneIntersectionVolume (4220-4307),
// boxParallelCutVolume (4313-4326), constructPolyhedronFaces (4333-4537),
// volumeOfPolyhedron (4548-4706), rotX / rotY (4719-4757) and
// volumeOfPyramid (4764-4785), FP32 in the reference's operation order.
// length(float2) is sqrt(x*x + y*y) and normalize(v) is v * rsqrt(dot(v, v))
// with rsqrt = 1 / sqrt: correctly rounded divide and square root.
// ---------------------------------------------------------------------------
__device__ __forceinline__ F3 fadd3(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 fscale3(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 rot_x(F3 v, float s, float c) {
    return {v.x, v.y * c - v.z * s, v.y * s + v.z * c};
}
__device__ __forceinline__ F3 rot_y(F3 v, float s, float c) {
    return {v.x * c + v.z * s, v.y, v.z * c - v.x * s};
}
__device__ __forceinline__ float len2f(float a, float b) { return sqrtf(a * a + b * b); }

// Turns the polygon and the apex so that rotN points along z, then returns
// the pyramid's volume (surveyor's formula times a third of the height).
__device__ __noinline__ float turned_pyramid(F3 *poly, int n, F3 rotN, F3 apex) {
    F3 h = apex;
    if (rotN.y != 0.0f) {
        const float pl = len2f(rotN.y, rotN.z), s = rotN.y / pl, c = rotN.z / pl;
        h = rot_x(h, s, c);
        for (int i = 0; i < n; i++) poly[i] = rot_x(poly[i], s, c);
        rotN = rot_x(rotN, s, c);
    }
    if (rotN.x != 0.0f) {
        const float pl = len2f(rotN.z, rotN.x), s = -rotN.x / pl, c = rotN.z / pl;
        h = rot_y(h, s, c);
        for (int i = 0; i < n; i++) poly[i] = rot_y(poly[i], s, c);
    }
    float area = 0.0f;
    for (int i = 0; i < n; i++) {
        const F3 a = poly[i], b = poly[(i + 1) % n];
        area += 0.5f * (a.x * b.y - b.x * a.y);
    }
    return fabsf((1.0f / 3.0f) * area * (h.z - poly[0].z));
}

// Volume the reference computes for the box centred at m + d * node, cut by
// the plane through t0 with normal n (the part the normal points into for
// a plane through the box).
__device__ __noinline__ float cut_volume(F3 t0, const int32_t node[3], F3 n, F3 m, float d) {
    const F3 vc = fadd3(m, fscale3(d, F3{float(node[0]), float(node[1]), float(node[2])}));
    if (fabsf(n.x) == 1.0f) return (d * d) * ((d / 2.0f) + t0.x - vc.x);
    if (fabsf(n.y) == 1.0f) return (d * d) * ((d / 2.0f) + t0.y - vc.y);
    if (fabsf(n.z) == 1.0f) return (d * d) * ((d / 2.0f) + t0.z - vc.z);
    const bool bx = n.x > 0.0f, by = n.y > 0.0f, bz = n.z > 0.0f;
    const float sx = float(!bx) * d + float(bx) * -d, sy = float(!by) * d + float(by) * -d,
                sz = float(!bz) * d + float(bz) * -d;
    const bool even = (int(bx) + int(by) + int(bz)) % 2 == 0;
    const F3 dx = {sx, 0.0f, 0.0f};
    const F3 dy = even ? F3{0.0f, sy, 0.0f} : F3{0.0f, 0.0f, sz};
    const F3 dz = even ? F3{0.0f, 0.0f, sz} : F3{0.0f, sy, 0.0f};
    const float dh = d / 2.0f;
    F3 v[8];
    v[0] = fadd3(vc, F3{dh * float(bx) - dh * float(!bx), dh * float(by) - dh * float(!by),
                        dh * float(bz) - dh * float(!bz)});
    v[1] = fadd3(v[0], dx);
    v[2] = fadd3(v[0], dy);
    v[3] = fadd3(v[0], dz);
    v[4] = fadd3(v[1], dz);
    v[5] = fadd3(v[2], dx);
    v[6] = fadd3(v[3], dy);
    v[7] = fadd3(v[4], dy);
    // Three paths from v0 (edges (0,a), (a,b), (b,7) and (a,c)) build the cut
    // polygon and faces 1-3; the arrays start zeroed (an empty cut polygon
    // still reads its first entry's z).
    F3 ip[6], face[3][6];
    for (int i = 0; i < 6; i++) {
        ip[i] = F3{0.0f, 0.0f, 0.0f};
        for (int f = 0; f < 3; f++) face[f][i] = F3{0.0f, 0.0f, 0.0f};
    }
    int ni = 0, nf[3] = {0, 0, 0};
    const int pa[3] = {1, 2, 3}, pb[3] = {4, 5, 6}, pc[3] = {5, 6, 4}, partner[3] = {2, 0, 1};
    for (int k = 0; k < 3; k++) {
        const int a = pa[k], b = pb[k], c = pc[k], f = k, g = partner[k];
        float t = fdot3(n, fsub(t0, v[0])) / fdot3(n, fsub(v[a], v[0]));
        if (t < 0.0f || t > 1.0f) {
            face[f][nf[f]++] = v[a];
            t = fdot3(n, fsub(t0, v[a])) / fdot3(n, fsub(v[b], v[a]));
            if (t < 0.0f || t > 1.0f) {
                t = fdot3(n, fsub(t0, v[b])) / fdot3(n, fsub(v[7], v[b]));
                if (t >= 0.0f && t <= 1.0f) {
                    ip[ni] = fadd3(v[b], fscale3(t, fsub(v[7], v[b])));
                    face[f][nf[f]++] = v[b];
                    face[f][nf[f]++] = ip[ni];
                    face[g][nf[g]++] = ip[ni];
                    face[g][nf[g]++] = v[b];
                    ni++;
                } else {
                    face[f][nf[f]++] = v[b];
                    face[g][nf[g]++] = v[b];
                }
            } else {
                ip[ni] = fadd3(v[a], fscale3(t, fsub(v[b], v[a])));
                face[f][nf[f]++] = ip[ni];
                ni++;
            }
        } else {
            ip[ni] = fadd3(v[0], fscale3(t, fsub(v[a], v[0])));
            ni++;
        }
        t = fdot3(n, fsub(t0, v[a])) / fdot3(n, fsub(v[c], v[a]));
        if (t >= 0.0f && t <= 1.0f) {
            ip[ni] = fadd3(v[a], fscale3(t, fsub(v[c], v[a])));
            face[f][nf[f]++] = ip[ni];
            ni++;
        }
    }
    float vol = 0.0f;
    vol += turned_pyramid(ip, ni, n, v[0]);
    for (int f = 0; f < 3; f++)
        if (nf[f] > 2) {
            const F3 e = fcross(fsub(face[f][0], face[f][2]), fsub(face[f][1], face[f][0]));
            const float inv = 1.0f / sqrtf(fdot3(e, e));
            vol += turned_pyramid(face[f], nf[f], F3{e.x * inv, e.y * inv, e.z * inv}, v[0]);
        }
    return vol;
}

// The ratio stored for the winner triangle t (class bb, 1 = 1D) at node
// coordinates `node` (the padded coordinates processVoxel receives): 1D
// triangles pass a zero normal; NaN is stored as 0x7FC00000.
__device__ __forceinline__ float partial_ratio(const float *__restrict__ V,
                                               const uint32_t *__restrict__ I, uint32_t t,
                                               uint32_t bb, const int32_t node[3], F3 m, float d) {
    const F3 t0 = fetch_f(V, I, t, 0), t1 = fetch_f(V, I, t, 1), t2 = fetch_f(V, I, t, 2);
    const F3 n = bb == 1u ? F3{0.0f, 0.0f, 0.0f} : fcross(fsub(t0, t2), fsub(t1, t0));
    const float vol = cut_volume(t0, node, n, m, d);
    const float r = fabsf(1.0f - vol / (d * d * d));
    return r != r ? __uint_as_float(0x7FC00000u) : r;
}

}  // namespace vk
