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

// classifyTriangles + analyzeBoundingBox: class (0..3) and dominant axis.
__device__ __forceinline__ void classify(const F3 t[3], F3 m, float d, int &bb, int &axis) {
    F3 n = fcross(fsub(t[0], t[2]), fsub(t[1], t[0]));
    if (!(n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)) {
        bb = BB_DEGEN;
        axis = AX_X;
        return;
    }
    F3 mn, mx;
    tri_bounds(t, mn, mx);
    VoxBB vb = voxel_bounds_half(mn, mx, m, d);
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

}  // namespace vk
