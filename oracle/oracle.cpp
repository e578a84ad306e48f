// oracle.cpp -- CPU restatement of hakarlss/Voxelizer's hot path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Written from the reference's
// algorithm description and source (cited per function); no reference code
// is copied.  Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
//
// Parity pins for each function live in tests/test_oracle_*.py.  Parity
// unpinned (no independent check is possible here): the rounding of the
// original nvcc build where it contracts to FMA, and which of several
// overlapping triangles' materials its racy store keeps (this restatement
// uses the deterministic rule of oracle.h).
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef uint32_t u32;

// ---------------------------------------------------------------------------
// Vector arithmetic in the operation order of the reference's vector helpers
// (include/helper_math.h: dot = a.x*b.x + a.y*b.y + a.z*b.z evaluated left to
// right, cross = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)).
// Written from those formulas; nothing of that file is reproduced.
// ---------------------------------------------------------------------------
struct D2 { double x, y; };
struct D3 { double x, y, z; };
struct F2 { float x, y; };
struct F3 { float x, y, z; };

inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float dot(F2 a, F2 b) { return a.x * b.x + a.y * b.y; }

// Saturating float -> int conversions: CUDA's cvt.rzi.{s,u}32.f{32,64} and
// gfx950's v_cvt_{i,u}32_f{32,64} both truncate toward zero, clamp to the
// destination range and map NaN to 0.
inline int32_t sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483648.0) return INT32_MAX;
    if (v <= -2147483649.0) return INT32_MIN;
    return (int32_t)v;
}
inline u32 sat_u32(double v) {
    if (v != v) return 0;
    if (v >= 4294967296.0) return UINT32_MAX;
    if (v <= -1.0) return 0;
    return (u32)v;
}
// 32-bit shifts with PTX semantics (shift amounts >= 32 give 0); the
// reference relies on them at src/voxelizer.cu:1579-1583,1607-1616.
inline u32 shr32(u32 v, u32 s) { return s >= 32u ? 0u : (v >> s); }
inline u32 shl32(u32 v, u32 s) { return s >= 32u ? 0u : (v << s); }

inline D3 fetch_d(const float *V, const u32 *I, u32 t, int k) {
    u32 i = I[3 * t + k];
    return {(double)V[3 * i], (double)V[3 * i + 1], (double)V[3 * i + 2]};
}
inline F3 fetch_f(const float *V, const u32 *I, u32 t, int k) {
    u32 i = I[3 * t + k];
    return {V[3 * i], V[3 * i + 1], V[3 * i + 2]};
}

// splitResolutionByNrOfParts for one axis (src/host_side.cpp:3003-3064): the
// length is divided from the top down with floor division.
struct Span { u32 min, max; };
std::vector<Span> split_axis(u32 len, u32 parts) {
    std::vector<Span> out(parts);
    u32 prev = len;
    for (u32 i = parts; i > 0; --i) {
        Span s;
        s.max = prev;
        s.min = s.max - s.max / i;
        prev = s.min;
        out[i - 1] = s;
    }
    return out;
}
// splitResolutionByMaxDim (src/host_side.cpp:2943-3002).
std::vector<Span> split_axis_max(u32 len, u32 maxdim) {
    u32 parts = (len + maxdim - 1) / maxdim;
    return split_axis(len, parts);
}

// ---------------------------------------------------------------------------
// Boundary-id table (src/voxelizer.cu:2142-2236; ids documented at
// include/voxelizer.h:47-83).  The table is an interface: the same permutation
// must give the same id, so the 27 (permutation, id) pairs are kept.
// Permutation bits: -x 1, +x 2, -y 4, +y 8, -z 16, +z 32.
// ---------------------------------------------------------------------------
const uint8_t kOrientPairs[27][2] = {
    {21, 1},  {22, 2},  {25, 3},  {26, 4},  {37, 5},  {38, 6},  {41, 7},
    {42, 8},  {23, 9},  {27, 10}, {29, 11}, {30, 12}, {39, 13}, {43, 14},
    {45, 15}, {46, 16}, {53, 17}, {54, 18}, {57, 19}, {58, 20}, {31, 21},
    {59, 22}, {55, 23}, {62, 24}, {61, 25}, {47, 26}, {63, 27}};

struct OrientTable {
    uint8_t lut[64];
    OrientTable() {
        std::memset(lut, 0, sizeof(lut));
        for (auto &p : kOrientPairs) lut[p[0]] = p[1];
    }
};
const OrientTable kOrient;

// ---------------------------------------------------------------------------
// Solid voxelization (src/voxelizer.cu:840-1617, src/host_side.cpp:275-397).
// ---------------------------------------------------------------------------
struct TileSetup {
    D2 ne[3];
    double de[3];
    u32 cmin[2], cmax[2];  // commonSpace (y, z), [min, max)
    u32 tmin[2], tmax[2];  // tile range (y, z), inclusive
};

// Shared prologue of calculateTileOverlap / constructWorkQueue
// (src/voxelizer.cu:861-945, 1080-1138).  Returns false where
// calculateTileOverlap writes 0 through one of its early outs (895-922).
bool tile_setup(const D3 v[3], const double extMin[3], double d, const u32 smin[2],
                const u32 smax[2], TileSetup &ts) {
    // getTriangleBounds(double3) (2398-2421): yz bounding box.
    D2 bmin = {v[0].y, v[0].z}, bmax = bmin;
    for (int i = 1; i < 3; i++) {
        bmin.x = v[i].y < bmin.x ? v[i].y : bmin.x;
        bmin.y = v[i].z < bmin.y ? v[i].z : bmin.y;
        bmax.x = v[i].y > bmax.x ? v[i].y : bmax.x;
        bmax.y = v[i].z > bmax.y ? v[i].z : bmax.y;
    }
    // getVoxelBoundsDouble (2450-2464).
    D2 vbmin = {(bmin.x - extMin[1]) / d, (bmin.y - extMin[2]) / d};
    D2 vbmax = {(bmax.x - extMin[1]) / d, (bmax.y - extMin[2]) / d};
    int32_t vmin[2] = {sat_i32(std::ceil(vbmin.x)), sat_i32(std::ceil(vbmin.y))};
    int32_t vmax[2] = {sat_i32(std::floor(vbmax.x) + 1.0), sat_i32(std::floor(vbmax.y) + 1.0)};
    // x component of the normal, in the reference's own expression (882-884).
    double tN = (v[0].y * v[1].z + v[2].y * v[0].z + v[2].z * v[1].y) -
                (v[2].y * v[1].z + v[0].z * v[1].y + v[2].z * v[0].y);
    if (tN == 0.0) return false;
    if (vmin[0] >= vmax[0] || vmin[1] >= vmax[1]) return false;
    if (vmax[0] <= (int32_t)smin[0] || vmax[1] <= (int32_t)smin[1]) return false;
    if (vmin[0] >= (int32_t)smax[0] || vmin[1] >= (int32_t)smax[1]) return false;
    for (int a = 0; a < 2; a++) {
        int32_t lo = std::max((int32_t)smin[a], vmin[a]);
        ts.cmin[a] = (u32)lo;
        ts.cmax[a] = std::min(smax[a], (u32)vmax[a]);
    }
    for (int i = 0; i < 3; i++) {
        const D3 &a = v[i], &b = v[(i + 1) % 3];
        if (tN > 0.0)
            ts.ne[i] = {a.z - b.z, b.y - a.y};
        else
            ts.ne[i] = {b.z - a.z, a.y - b.y};
    }
    for (int i = 0; i < 3; i++) ts.de[i] = -(ts.ne[i].x * v[i].y + ts.ne[i].y * v[i].z);
    for (int a = 0; a < 2; a++) {
        ts.tmin[a] = ts.cmin[a] / 4;
        ts.tmax[a] = (ts.cmax[a] - 1) / 4;
    }
    return true;
}

// traverseTilesForOverlaps / traverseTilesForWorkQueue (980-1033, 1169-1232):
// a tile is queued if any voxel centre of it inside commonSpace passes the
// (>= 0) edge tests.
template <class F>
void traverse_tiles(const TileSetup &ts, const double extMin[3], double d, F &&emit) {
    for (u32 t = ts.tmin[0]; t <= ts.tmax[0]; t++) {
        for (u32 u = ts.tmin[1]; u <= ts.tmax[1]; u++) {
            bool hit = false;
            for (u32 i = t * 4; i < t * 4 + 4 && i < ts.cmax[0] && !hit; i++) {
                for (u32 j = u * 4; j < u * 4 + 4 && j < ts.cmax[1] && !hit; j++) {
                    if (i >= ts.cmin[0] && j >= ts.cmin[1]) {
                        D2 p = {double(i) * d + extMin[1], double(j) * d + extMin[2]};
                        bool r = true;
                        for (int k = 0; k < 3; k++)
                            r &= (ts.ne[k].x * p.x + ts.ne[k].y * p.y + ts.de[k] >= 0.0);
                        if (r) hit = true;
                    }
                }
            }
            if (hit) emit(t, u);
        }
    }
}

// Per-triangle data of generateVoxelization (1334-1372).
struct FillSetup {
    D3 v0, n;
    D2 ne[3];
    double de[3];
};
FillSetup fill_setup(const D3 v[3]) {
    FillSetup fs;
    fs.v0 = v[0];
    fs.n = cross(sub(v[0], v[2]), sub(v[1], v[0]));
    for (int i = 0; i < 3; i++) {
        const D3 &a = v[i], &b = v[(i + 1) % 3];
        if (fs.n.x > 0.0)
            fs.ne[i] = {a.z - b.z, b.y - a.y};
        else
            fs.ne[i] = {b.z - a.z, a.y - b.y};
        fs.de[i] = -(fs.ne[i].x * a.y + fs.ne[i].y * a.z);
    }
    return fs;
}
// Centre test + crossing (1418-1453): returns the x voxel (incl. +1 padding).
bool fill_center(const FillSetup &fs, const double extMin[3], double d, u32 cy, u32 cz,
                 u32 &voxelX) {
    D2 p = {double(cy) * d + extMin[1], double(cz) * d + extMin[2]};
    u32 ok = 0;
    for (int i = 0; i < 3; i++) {
        double r = fs.ne[i].x * p.x + fs.ne[i].y * p.y + fs.de[i];
        if (fs.ne[i].x > 0.0 || (fs.ne[i].x == 0.0 && fs.ne[i].y < 0.0)) r += DBL_MIN;
        if (r > 0.0) ok++;
    }
    if (ok != 3) return false;
    D3 A = {fs.v0.x - extMin[0], fs.v0.y - p.x, fs.v0.z - p.y};
    double B = A.x * fs.n.x + A.y * fs.n.y + A.z * fs.n.z;
    double px = B / fs.n.x;
    px = px < 0.0 ? 0.0 : px;
    voxelX = sat_u32(std::ceil(px / d));
    voxelX += 1u;
    return true;
}

struct SolidCtx {
    const float *V;
    const u32 *I;
    u32 nt;
    double extMin[3];
    double d;
    u32 X, Yext, Zext;  // ext resolution (x already padded/rounded)
    u32 xSplits;
    u32 *bits;
    // Allocated rows and planes, and the alloc index of ext row / plane 0:
    // 1 (padding before it), or 0 where another subspace lies before the
    // partition and ext row 0 is its overlap row (src/voxelizer.cu:1487-1502).
    u32 Yalloc, Zalloc, oy = 1, oz = 1;
};

// generateVoxelization for one tile and one x split, per voxel-centre lane
// (1273-1617), with the reference's active segment / word bitmask bookkeeping
// and its atomicXor flushes applied to the bit array.
void fill_tile_lane(const SolidCtx &c, const std::vector<FillSetup> &tris, u32 tileNr,
                    const u32 sub_min[3], const u32 sub_max[3], int lane) {
    const u32 tilesY = (c.Yext + 3) / 4;
    const u32 tileZ = tileNr / tilesY, tileY = tileNr % tilesY;
    const u32 vZ0 = u32(lane / 4) + tileZ * 4;
    const u32 vY0 = u32(lane % 4) + tileY * 4;
    if (!(vY0 >= sub_min[1] && vY0 < sub_max[1] && vZ0 >= sub_min[2] && vZ0 < sub_max[2]))
        return;
    u32 active = 0, current = UINT32_MAX, bitmask = 0;
    const u32 resY = c.Yalloc, resXw = c.X >> 5;
    const u32 vY = vY0 + c.oy, vZ = vZ0 + c.oz;
    for (const FillSetup &fs : tris) {
        u32 voxelX;
        if (!fill_center(fs, c.extMin, c.d, vY0, vZ0, voxelX)) continue;
        u32 ix = voxelX >> 5;
        u32 min_ix = (sub_min[0] + 2) >> 5;
        if (ix >= ((sub_max[0] + 2) >> 5)) {
            // crossing beyond this x split: ignored
        } else if (ix < min_ix) {
            u32 inti = resY * resXw * vZ + resXw * vY + min_ix;
            if (current != inti) {
                if (active != 0) c.bits[current] ^= active;
                active = 0;
                current = inti;
            }
            bitmask ^= UINT32_MAX;
        } else {
            u32 inti = resY * resXw * vZ + resXw * vY + ix;
            if (current != inti) {
                if (active != 0) c.bits[current] ^= active;
                active = 0;
                current = inti;
            }
            active ^= UINT32_MAX >> (voxelX % 32u);
            u32 lastIX = (sub_max[0] + 1) >> 5;
            u32 ints = (sub_max[0] - sub_min[0]) >> 5;
            u32 bm = ix < lastIX ? shr32(UINT32_MAX, ix % ints + 1u) : 0u;
            bitmask ^= bm;
        }
    }
    if (current == UINT32_MAX) return;
    if (active != 0) c.bits[current] ^= active;
    u32 intsPerDim = (sub_max[0] - sub_min[0]) >> 5;
    u32 start = current - (current % intsPerDim);
    for (u32 i = 0; i < intsPerDim; i++) {
        u32 value = (bitmask & shl32(1u, 31u - i)) > 0 ? UINT32_MAX : 0u;
        if (value != 0) c.bits[start + i] ^= value;
    }
}

// performVoxelization for one y/z subspace (src/host_side.cpp:275-397).
void perform_voxelization(const SolidCtx &c, Span ys, Span zs) {
    const u32 smin[2] = {ys.min, zs.min}, smax[2] = {ys.max, zs.max};
    // calculateTileOverlap + prepareForConstructWorkQueue + constructWorkQueue:
    // the queue holds (tile, triangle) pairs; sortWorkQueue/compactWorkQueue
    // group it by tile.  Grouping order does not matter (XOR commutes).
    std::vector<std::pair<u32, u32>> queue;  // (tileId, triangle)
    const u32 tilesY = (c.Yext + 3) / 4;
    for (u32 t = 0; t < c.nt; t++) {
        D3 v[3] = {fetch_d(c.V, c.I, t, 0), fetch_d(c.V, c.I, t, 1), fetch_d(c.V, c.I, t, 2)};
        TileSetup ts;
        if (!tile_setup(v, c.extMin, c.d, smin, smax, ts)) continue;
        traverse_tiles(ts, c.extMin, c.d,
                       [&](u32 ty, u32 tz) { queue.push_back({tz * tilesY + ty, t}); });
    }
    if (queue.empty()) return;
    std::stable_sort(queue.begin(), queue.end(),
                     [](const std::pair<u32, u32> &a, const std::pair<u32, u32> &b) {
                         return a.first < b.first;
                     });
    std::vector<size_t> starts;
    for (size_t i = 0; i < queue.size(); i++)
        if (i == 0 || queue[i].first != queue[i - 1].first) starts.push_back(i);
    starts.push_back(queue.size());
    const u32 xRes = c.X / c.xSplits;
    for (u32 xs = 0; xs < c.xSplits; xs++) {
        const u32 sub_min[3] = {xs * xRes, ys.min, zs.min};
        const u32 sub_max[3] = {(xs + 1) * xRes, ys.max, zs.max};
        const long ntiles = (long)starts.size() - 1;
#pragma omp parallel for schedule(dynamic, 16)
        for (long g = 0; g < ntiles; g++) {
            std::vector<FillSetup> tris;
            for (size_t k = starts[g]; k < starts[g + 1]; k++) {
                u32 t = queue[k].second;
                D3 v[3] = {fetch_d(c.V, c.I, t, 0), fetch_d(c.V, c.I, t, 1),
                           fetch_d(c.V, c.I, t, 2)};
                tris.push_back(fill_setup(v));
            }
            // Each lane owns one row, so tiles never share bit words.
            for (int lane = 0; lane < 16; lane++)
                fill_tile_lane(c, tris, queue[starts[g]].first, sub_min, sub_max, lane);
        }
    }
}

// Ext space and alloc layout of one partition (one device's subspace of
// voxelizeWorker, src/host_side.cpp:901-1011); the whole grid is the
// partition with no neighbours.
struct PartLayout {
    u32 Yext, Zext;      // ext rows / planes (partition + overlap rows)
    u32 Yalloc, Zalloc;  // allocated rows / planes (partition + 2)
    u32 oy, oz;          // alloc index of ext row / plane 0
    double extMin[3];    // minVertex + extPartition.min * d
    bool any_neighbour;
};
PartLayout whole_layout(const vox_grid &g, const double extMin[3]) {
    PartLayout L;
    L.Yext = g.res[1];
    L.Zext = g.res[2];
    L.Yalloc = g.alloc[1];
    L.Zalloc = g.alloc[2];
    L.oy = L.oz = 1;
    for (int a = 0; a < 3; a++) L.extMin[a] = extMin[a];
    L.any_neighbour = false;
    return L;
}

void solid_bits_layout(const float *V, const u32 *I, u32 nt, const vox_grid &g,
                       const PartLayout &L, u32 splitY, u32 splitZ, u32 *bits) {
    SolidCtx c;
    c.V = V;
    c.I = I;
    c.nt = nt;
    for (int a = 0; a < 3; a++) c.extMin[a] = L.extMin[a];
    c.d = g.voxel_length;
    c.X = g.res[0];
    c.Yext = L.Yext;
    c.Zext = L.Zext;
    c.xSplits = g.x_splits;
    c.bits = bits;
    c.Yalloc = L.Yalloc;
    c.Zalloc = L.Zalloc;
    c.oy = L.oy;
    c.oz = L.oz;
    size_t words = (size_t)g.alloc[0] / 32 * L.Yalloc * L.Zalloc;
    std::memset(bits, 0, words * 4);
    // extYzSplits over extPartition; subspaces ordered z outer, y inner
    // (src/host_side.cpp:3100-3112).
    std::vector<Span> ys = split_axis_max(c.Yext, splitY), zs = split_axis_max(c.Zext, splitZ);
    for (const Span &z : zs)
        for (const Span &y : ys) perform_voxelization(c, y, z);
}

void solid_bits(const float *V, const u32 *I, u32 nt, const vox_grid &g,
                const double extMin[3], u32 splitY, u32 splitZ, u32 *bits) {
    solid_bits_layout(V, I, nt, g, whole_layout(g, extMin), splitY, splitZ, bits);
}

// ---------------------------------------------------------------------------
// Surface voxelization (src/voxelizer.cu:2364-2590, 2677-2757, 2892-2942,
// 3040-3221, 3307-3993, 4004-4201).  FP32 throughout.
// ---------------------------------------------------------------------------
enum { BB_DEGEN = 0, BB_1D = 1, BB_2D = 2, BB_3D = 3 };
enum { AX_X = 0, AX_Y = 1, AX_Z = 2 };

struct I3 { int32_t x, y, z; };
struct VoxBB { I3 min, max; };

struct OverlapData { F2 ne[3]; float de[3]; };

inline int32_t f2i(float v) { return sat_i32((double)v); }

// getTriangleBounds(float3) (2364-2392): strict comparisons.
void tri_bounds(const F3 t[3], F3 &mn, F3 &mx) {
    mn = t[0];
    mx = t[0];
    for (int i = 1; i < 3; i++) {
        if (t[i].x < mn.x) mn.x = t[i].x;
        if (t[i].y < mn.y) mn.y = t[i].y;
        if (t[i].z < mn.z) mn.z = t[i].z;
        if (t[i].x > mx.x) mx.x = t[i].x;
        if (t[i].y > mx.y) mx.y = t[i].y;
        if (t[i].z > mx.z) mx.z = t[i].z;
    }
}
// getVoxelBoundsHalf (2471-2492).
VoxBB voxel_bounds_half(F3 mn, F3 mx, F3 m, float d) {
    VoxBB b;
    b.min = {f2i(std::ceil((mn.x - m.x) / d - 0.5f)), f2i(std::ceil((mn.y - m.y) / d - 0.5f)),
             f2i(std::ceil((mn.z - m.z) / d - 0.5f))};
    b.max = {f2i(std::floor((mx.x - m.x) / d + 0.5f)), f2i(std::floor((mx.y - m.y) / d + 0.5f)),
             f2i(std::floor((mx.z - m.z) / d + 0.5f))};
    return b;
}
inline int32_t vox_cbound_min(float v, float m, float d) {  // 2499-2510
    return f2i(std::ceil((v - m) / d - 0.5f));
}
inline int32_t vox_cbound_max(float v, float m, float d) {  // 2513-2524
    return f2i(std::floor((v - m) / d + 0.5f));
}
// determineDominantAxis (2531-2550).
int dominant_axis(F3 n) {
    float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
    if (ax > ay && ax > az) return AX_X;
    if (ay > ax && ay > az) return AX_Y;
    return AX_Z;
}
// getSingleVoxelBoundsComponent (2578-2590).
inline float vox_corner(int32_t c, float m, float d) { return m + (float(c) - 0.5f) * d; }

// analyzeBoundingBox (3117-3205) + classifyTriangles (3040-3105).
struct TriClass { int bb, axis; u32 cols; };
TriClass classify(const F3 t[3], F3 m, float d) {
    F3 n = cross(sub(t[0], t[2]), sub(t[1], t[0]));
    TriClass tc;
    if (!(n.x != 0.0f || n.y != 0.0f || n.z != 0.0f)) {
        tc.bb = BB_DEGEN;
        tc.axis = AX_X;
        tc.cols = 0;
        return tc;
    }
    F3 mn, mx;
    tri_bounds(t, mn, mx);
    VoxBB vb = voxel_bounds_half(mn, mx, m, d);
    I3 df = {vb.max.x - vb.min.x, vb.max.y - vb.min.y, vb.max.z - vb.min.z};
    if (df.x == 0) {
        if (df.y == 0) {
            if (df.z == 0) tc = {BB_1D, AX_X, 1u};
            else tc = {BB_1D, AX_Z, (u32)df.z};
        } else if (df.z == 0) {
            tc = {BB_1D, AX_Y, (u32)df.y};
        } else {
            tc = {BB_2D, AX_X, (u32)(df.y * df.z)};
        }
    } else if (df.y == 0) {
        if (df.z == 0) tc = {BB_1D, AX_X, (u32)df.x};
        else tc = {BB_2D, AX_Y, (u32)(df.x * df.z)};
    } else if (df.z == 0) {
        tc = {BB_2D, AX_Z, (u32)(df.x * df.y)};
    } else {
        tc.bb = BB_3D;
        tc.axis = dominant_axis(n);
        if (tc.axis == AX_X) tc.cols = (u32)(df.y * df.z);
        else if (tc.axis == AX_Y) tc.cols = (u32)(df.x * df.z);
        else tc.cols = (u32)(df.x * df.y);
    }
    return tc;
}
// encodeTriangleType (3214-3221).
inline u32 encode_type(const TriClass &c) {
    return c.cols | ((u32)c.axis << 28) | ((u32)c.bb << 30);
}

// setup{YZ,ZX,XY}OverlapTest (2677-2757).
OverlapData setup_yz(const F3 t[3], F3 n, float d) {
    OverlapData o;
    for (int i = 0; i < 3; i++) {
        const F3 &a = t[i], &b = t[(i + 1) % 3];
        if (n.x >= 0.0f) o.ne[i] = {a.z - b.z, b.y - a.y};
        else o.ne[i] = {b.z - a.z, a.y - b.y};
        o.de[i] = -(o.ne[i].x * a.y + o.ne[i].y * a.z) + std::fmax(0.0f, d * o.ne[i].x) +
                  std::fmax(0.0f, d * o.ne[i].y);
    }
    return o;
}
OverlapData setup_zx(const F3 t[3], F3 n, float d) {
    OverlapData o;
    for (int i = 0; i < 3; i++) {
        const F3 &a = t[i], &b = t[(i + 1) % 3];
        if (n.y >= 0.0f) o.ne[i] = {a.x - b.x, b.z - a.z};
        else o.ne[i] = {b.x - a.x, a.z - b.z};
        o.de[i] = -(o.ne[i].x * a.z + o.ne[i].y * a.x) + std::fmax(0.0f, d * o.ne[i].x) +
                  std::fmax(0.0f, d * o.ne[i].y);
    }
    return o;
}
OverlapData setup_xy(const F3 t[3], F3 n, float d) {
    OverlapData o;
    for (int i = 0; i < 3; i++) {
        const F3 &a = t[i], &b = t[(i + 1) % 3];
        if (n.z >= 0.0f) o.ne[i] = {a.y - b.y, b.x - a.x};
        else o.ne[i] = {b.y - a.y, a.x - b.x};
        o.de[i] = -(o.ne[i].x * a.x + o.ne[i].y * a.y) + std::fmax(0.0f, d * o.ne[i].x) +
                  std::fmax(0.0f, d * o.ne[i].y);
    }
    return o;
}
// overlapTest{YZ,ZX,XY} (2892-2942).
inline bool test2(const OverlapData &o, F2 p2) {
    bool r = true;
    for (int i = 0; i < 3; i++) r &= dot(o.ne[i], p2) + o.de[i] >= 0.0f;
    return r;
}
inline bool test_yz(const OverlapData &o, F3 p) { return test2(o, {p.y, p.z}); }
inline bool test_zx(const OverlapData &o, F3 p) { return test2(o, {p.z, p.x}); }
inline bool test_xy(const OverlapData &o, F3 p) { return test2(o, {p.x, p.y}); }

// intersectWithPlane{X,Y,Z} (4147-4182) and determineDepthRange{X,Y,Z}
// (4004-4119).
inline float isect(F3 p, F3 v, F3 n, int axis) {
    float num = dot(sub(v, p), n);
    return num / (axis == AX_X ? n.x : axis == AX_Y ? n.y : n.z);
}
inline void bounds_check(float &a, float &b) {
    if (a > b) {
        float t = a;
        a = b;
        b = t;
    }
}
void depth_range(const F3 t[3], F3 p, F3 n, F3 m, float d, int axis, int32_t &lo,
                 int32_t &hi) {
    F3 pmin, pmax;
    float a, b;
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
    a = isect(pmin, t[0], n, axis);
    b = isect(pmax, t[0], n, axis);
    bounds_check(a, b);
    float mc = axis == AX_X ? m.x : axis == AX_Y ? m.y : m.z;
    lo = vox_cbound_min(a, mc, d);
    hi = vox_cbound_max(b, mc, d);
}

// adjustVoxelRange (3963-3993); returns false where it marks min = -1.
bool adjust_range(VoxBB &vb, const u32 smin[3], const u32 smax[3]) {
    I3 lo = {(int32_t)smin[0], (int32_t)smin[1], (int32_t)smin[2]};
    I3 hi = {(int32_t)smax[0] - 1, (int32_t)smax[1] - 1, (int32_t)smax[2] - 1};
    if (lo.x > vb.max.x || lo.y > vb.max.y || lo.z > vb.max.z) return false;
    if (vb.min.x > hi.x || vb.min.y > hi.y || vb.min.z > hi.z) return false;
    vb.min = {std::max(vb.min.x, lo.x), std::max(vb.min.y, lo.y), std::max(vb.min.z, lo.z)};
    vb.max = {std::min(vb.max.x, hi.x), std::min(vb.max.y, hi.y), std::min(vb.max.z, hi.z)};
    return true;
}

inline void atomic_max_u32(u32 *p, u32 v) {
    u32 cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
    }
}

// Where processVoxel (2946-2989) stores: node index of the (already +1
// adjusted) voxel for a Cartesian grid or FCC sub-grid g.
struct SurfTarget {
    u32 resX, resY, resZ;  // allocated per-grid dims (ext dims + padding)
    int gridType;          // 0: Cartesian, 1..4: FCC sub-grid
    u32 *keys;
    size_t nkeys;
    // Cartesian partitions: the writes' y / z adjustment is oy / oz, not 1
    // (src/voxelizer.cu:3374-3383).
    int32_t dy = 0, dz = 0;
    size_t index(int32_t x, int32_t y, int32_t z) const {
        if (gridType == 0) {
            y += dy;
            z += dz;
            return (size_t)resX * resY * (u32)z + (size_t)resX * (u32)y + (u32)x;
        }
        size_t l = 2 * (size_t)resX, a = l * resY;
        size_t ox = gridType == 2 || gridType == 4, oz = gridType == 3 || gridType == 4;
        return a * (2 * (size_t)(u32)z + oz) + l * (u32)y + 2 * (size_t)(u32)x + ox;
    }
    void write(int32_t x, int32_t y, int32_t z, u32 key) const {
        size_t i = index(x, y, z);
        if (i < nkeys) atomic_max_u32(&keys[i], key);
    }
};

// process{1D,2D,3D}Triangles for one triangle and one subspace
// (3307-3411, 3419-3596, 3604-3955).  Coordinates are ext-relative; the
// single-device padding adjustment is (1, 1, 1).
void surface_triangle(const F3 t[3], u32 triIdx, u32 type, F3 m, float d, const u32 smin[3],
                      const u32 smax[3], const SurfTarget &st) {
    const int bb = (int)(type >> 30);
    const int axis = (int)((type >> 28) & 3u);
    const u32 key = ((u32)bb << 30) | triIdx;
    F3 mn, mx;
    tri_bounds(t, mn, mx);
    VoxBB vb = voxel_bounds_half(mn, mx, m, d);
    if (!adjust_range(vb, smin, smax)) return;
    if (bb == BB_1D) {
        I3 lo = {vb.min.x + 1, vb.min.y + 1, vb.min.z + 1};
        I3 hi = {vb.max.x + 1, vb.max.y + 1, vb.max.z + 1};
        if (axis == AX_X)
            for (int32_t i = lo.x; i <= hi.x; i++) st.write(i, lo.y, lo.z, key);
        else if (axis == AX_Y)
            for (int32_t j = lo.y; j <= hi.y; j++) st.write(lo.x, j, lo.z, key);
        else
            for (int32_t k = lo.z; k <= hi.z; k++) st.write(lo.x, lo.y, k, key);
        return;
    }
    F3 n = cross(sub(t[0], t[2]), sub(t[1], t[0]));
    if (bb == BB_2D) {
        if (axis == AX_X) {
            OverlapData o = setup_yz(t, n, d);
            for (int32_t j = vb.min.y; j <= vb.max.y; j++)
                for (int32_t k = vb.min.z; k <= vb.max.z; k++) {
                    F3 p = {0.0f, vox_corner(j, m.y, d), vox_corner(k, m.z, d)};
                    if (test_yz(o, p)) st.write(vb.min.x + 1, j + 1, k + 1, key);
                }
        } else if (axis == AX_Y) {
            OverlapData o = setup_zx(t, n, d);
            for (int32_t i = vb.min.x; i <= vb.max.x; i++)
                for (int32_t k = vb.min.z; k <= vb.max.z; k++) {
                    F3 p = {vox_corner(i, m.x, d), 0.0f, vox_corner(k, m.z, d)};
                    if (test_zx(o, p)) st.write(i + 1, vb.min.y + 1, k + 1, key);
                }
        } else {
            OverlapData o = setup_xy(t, n, d);
            for (int32_t i = vb.min.x; i <= vb.max.x; i++)
                for (int32_t j = vb.min.y; j <= vb.max.y; j++) {
                    F3 p = {vox_corner(i, m.x, d), vox_corner(j, m.y, d), 0.0f};
                    if (test_xy(o, p)) st.write(i + 1, j + 1, vb.min.z + 1, key);
                }
        }
        return;
    }
    if (bb != BB_3D) return;
    OverlapData oyz = setup_yz(t, n, d), ozx = setup_zx(t, n, d), oxy = setup_xy(t, n, d);
    if (axis == AX_X) {
        for (int32_t j = vb.min.y; j <= vb.max.y; j++)
            for (int32_t k = vb.min.z; k <= vb.max.z; k++) {
                F3 p = {0.0f, vox_corner(j, m.y, d), vox_corner(k, m.z, d)};
                if (!test_yz(oyz, p)) continue;
                int32_t lo, hi;
                depth_range(t, p, n, m, d, AX_X, lo, hi);
                if (lo > vb.max.x || hi < vb.min.x) continue;
                lo = lo < vb.min.x ? vb.min.x : lo;
                hi = hi > vb.max.x ? vb.max.x : hi;
                if (lo == hi) {
                    st.write(lo + 1, j + 1, k + 1, key);
                } else {
                    for (int32_t i = lo; i <= hi; i++) {
                        p.x = vox_corner(i, m.x, d);
                        if (test_zx(ozx, p) && test_xy(oxy, p)) st.write(i + 1, j + 1, k + 1, key);
                    }
                }
            }
    } else if (axis == AX_Y) {
        for (int32_t k = vb.min.z; k <= vb.max.z; k++)
            for (int32_t i = vb.min.x; i <= vb.max.x; i++) {
                F3 p = {vox_corner(i, m.x, d), 0.0f, vox_corner(k, m.z, d)};
                if (!test_zx(ozx, p)) continue;
                int32_t lo, hi;
                depth_range(t, p, n, m, d, AX_Y, lo, hi);
                if (lo > vb.max.y || hi < vb.min.y) continue;
                lo = lo < vb.min.y ? vb.min.y : lo;
                hi = hi > vb.max.y ? vb.max.y : hi;
                if (lo == hi) {
                    st.write(i + 1, lo + 1, k + 1, key);
                } else {
                    for (int32_t j = lo; j <= hi; j++) {
                        p.y = vox_corner(j, m.y, d);
                        if (test_xy(oxy, p) && test_yz(oyz, p)) st.write(i + 1, j + 1, k + 1, key);
                    }
                }
            }
    } else {
        for (int32_t i = vb.min.x; i <= vb.max.x; i++)
            for (int32_t j = vb.min.y; j <= vb.max.y; j++) {
                F3 p = {vox_corner(i, m.x, d), vox_corner(j, m.y, d), 0.0f};
                if (!test_xy(oxy, p)) continue;
                int32_t lo, hi;
                depth_range(t, p, n, m, d, AX_Z, lo, hi);
                if (lo > vb.max.z || hi < vb.min.z) continue;
                lo = lo < vb.min.z ? vb.min.z : lo;
                hi = hi > vb.max.z ? vb.max.z : hi;
                if (lo == hi) {
                    st.write(i + 1, j + 1, lo + 1, key);
                } else {
                    for (int32_t k = lo; k <= hi; k++) {
                        p.z = vox_corner(k, m.z, d);
                        if (test_yz(oyz, p) && test_zx(ozx, p)) st.write(i + 1, j + 1, k + 1, key);
                    }
                }
            }
    }
}

// calcTriangleClassification + calcOptSurfaceVoxelization for one grid
// (src/voxelizer.cu:505-724; splits as in src/host_side.cpp:1107-1134).
void surface_grid_res(const float *V, const u32 *I, u32 nt, const vox_grid &g,
                      const double extMin[3], const u32 extRes[3], const u32 matSplit[3],
                      const SurfTarget &st);
void surface_grid(const float *V, const u32 *I, u32 nt, const vox_grid &g,
                  const double extMin[3], const u32 matSplit[3], const SurfTarget &st) {
    surface_grid_res(V, I, nt, g, extMin, g.res, matSplit, st);
}
// The surface stage over an ext space of extRes voxels (matSplitRes splits
// the ext partition, src/host_side.cpp:1107-1111).
void surface_grid_res(const float *V, const u32 *I, u32 nt, const vox_grid &g,
                      const double extMin[3], const u32 extRes[3], const u32 matSplit[3],
                      const SurfTarget &st) {
    const F3 m = {(float)extMin[0], (float)extMin[1], (float)extMin[2]};
    const float d = (float)g.voxel_length;
    std::vector<u32> types(nt);
    for (u32 t = 0; t < nt; t++) {
        F3 tri[3] = {fetch_f(V, I, t, 0), fetch_f(V, I, t, 1), fetch_f(V, I, t, 2)};
        types[t] = encode_type(classify(tri, m, d));
    }
    // The reference sorts by type and runs the 1D, 2D and 3D kernels on the
    // class ranges; with the max-key winner rule only membership matters.
    std::vector<Span> xs = split_axis_max(extRes[0], matSplit[0]);
    std::vector<Span> ys = split_axis_max(extRes[1], matSplit[1]);
    std::vector<Span> zs = split_axis_max(extRes[2], matSplit[2]);
    for (const Span &z : zs)
        for (const Span &y : ys)
            for (const Span &x : xs) {
                const u32 smin[3] = {x.min, y.min, z.min}, smax[3] = {x.max, y.max, z.max};
#pragma omp parallel for schedule(dynamic, 256)
                for (long t = 0; t < (long)nt; t++) {
                    u32 bb = types[t] >> 30;
                    if (bb == BB_DEGEN) continue;
                    F3 tri[3] = {fetch_f(V, I, (u32)t, 0), fetch_f(V, I, (u32)t, 1),
                                 fetch_f(V, I, (u32)t, 2)};
                    surface_triangle(tri, (u32)t, types[t], m, d, smin, smax, st);
                }
            }
}

// ---------------------------------------------------------------------------
// PartialNode volume ratio: processVoxel<PartialNode> (src/voxelizer.cu:
// 2997-3033) with calculateVoxelPlaneIntersectionVolume (4220-4307),
// boxParallelCutVolume (4313-4326), constructPolyhedronFaces (4333-4537),
// volumeOfPolyhedron (4548-4706), rotX / rotY (4719-4757) and
// volumeOfPyramid (4764-4785).  FP32 in the reference's operation order.
// length(float2 v) is sqrtf(v.x*v.x + v.y*v.y) and normalize(v) is
// v * rsqrtf(dot(v, v)); rsqrtf is taken as 1.0f / sqrtf(x), two correctly
// rounded operations (CUDA's approximate rsqrtf is not reproduced).
// ---------------------------------------------------------------------------
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 scale(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline F3 rot_x(F3 v, float s, float c) { return {v.x, v.y * c - v.z * s, v.y * s + v.z * c}; }
inline F3 rot_y(F3 v, float s, float c) { return {v.x * c + v.z * s, v.y, v.z * c - v.x * s}; }
inline float len2(float a, float b) { return std::sqrt(a * a + b * b); }
inline F3 normalize3(F3 v) {
    const float inv = 1.0f / std::sqrt(dot(v, v));
    return {v.x * inv, v.y * inv, v.z * inv};
}

// Surveyor's formula over the base, times a third of the height.
float pyramid_volume(const F3 *base, int n, F3 h) {
    float area = 0.0f;
    for (int i = 0; i < n; i++) {
        const F3 a = base[i], b = base[(i + 1) % n];
        area += 0.5f * (a.x * b.y - b.x * a.y);
    }
    return std::fabs((1.0f / 3.0f) * area * (h.z - base[0].z));
}

// Turns the polygon (and the apex) so that its normal n points along z, then
// adds its pyramid (the three face polygons pass their own normal).
float turned_pyramid(F3 *poly, int n, F3 rotN, F3 apex) {
    F3 h = apex;
    if (rotN.y != 0.0f) {
        const float pl = len2(rotN.y, rotN.z), s = rotN.y / pl, c = rotN.z / pl;
        h = rot_x(h, s, c);
        for (int i = 0; i < n; i++) poly[i] = rot_x(poly[i], s, c);
        rotN = rot_x(rotN, s, c);
    }
    if (rotN.x != 0.0f) {
        const float pl = len2(rotN.z, rotN.x), s = -rotN.x / pl, c = rotN.z / pl;
        h = rot_y(h, s, c);
        for (int i = 0; i < n; i++) poly[i] = rot_y(poly[i], s, c);
    }
    return pyramid_volume(poly, n, h);
}

// Volume of the part of the voxel at `voxel` (centre m + d * voxel) behind
// the plane of the triangle with normal n: the reference's exact arithmetic,
// including its results for normals that do not cut the box.
float cut_volume(const F3 tri[3], const int32_t voxel[3], F3 n, F3 m, float d) {
    const F3 vc = add(m, scale(d, F3{float(voxel[0]), float(voxel[1]), float(voxel[2])}));
    if (std::fabs(n.x) == 1.0f) return (d * d) * ((d / 2.0f) + tri[0].x - vc.x);
    if (std::fabs(n.y) == 1.0f) return (d * d) * ((d / 2.0f) + tri[0].y - vc.y);
    if (std::fabs(n.z) == 1.0f) return (d * d) * ((d / 2.0f) + tri[0].z - vc.z);
    const bool bx = n.x > 0.0f, by = n.y > 0.0f, bz = n.z > 0.0f;
    auto step = [d](bool b) { return float(!b) * d + float(b) * -d; };
    F3 dx = {step(bx), 0.0f, 0.0f}, dy, dz;
    if ((int(bx) + int(by) + int(bz)) % 2 == 0) {
        dy = {0.0f, step(by), 0.0f};
        dz = {0.0f, 0.0f, step(bz)};
    } else {
        dz = {0.0f, step(by), 0.0f};
        dy = {0.0f, 0.0f, step(bz)};
    }
    const float dh = d / 2.0f;
    auto half = [dh](bool b) { return dh * float(b) - dh * float(!b); };
    F3 v[8];
    v[0] = add(vc, F3{half(bx), half(by), half(bz)});
    v[1] = add(v[0], dx);
    v[2] = add(v[0], dy);
    v[3] = add(v[0], dz);
    v[4] = add(v[1], dz);
    v[5] = add(v[2], dx);
    v[6] = add(v[3], dy);
    v[7] = add(v[4], dy);
    // constructPolyhedronFaces: three paths from v0, each edge tested for the
    // plane; the arrays start zeroed (an empty cut polygon still reads its
    // first entry's z in volumeOfPyramid).
    F3 ip[6] = {}, face[3][6] = {};
    int ni = 0, nf[3] = {0, 0, 0};
    auto T = [&](int p, int q) { return dot(n, sub(tri[0], v[p])) / dot(n, sub(v[q], v[p])); };
    auto at = [&](int p, int q, float t) { return add(v[p], scale(t, sub(v[q], v[p]))); };
    // path k: edges (0,a), (a,b), (b,7) and (a,c); face k, partner face k-1
    const int pa[3] = {1, 2, 3}, pb[3] = {4, 5, 6}, pc[3] = {5, 6, 4}, partner[3] = {2, 0, 1};
    for (int k = 0; k < 3; k++) {
        const int a = pa[k], b = pb[k], c = pc[k], f = k, g = partner[k];
        float t = T(0, a);
        if (t < 0.0f || t > 1.0f) {
            face[f][nf[f]++] = v[a];
            t = T(a, b);
            if (t < 0.0f || t > 1.0f) {
                t = T(b, 7);
                if (t >= 0.0f && t <= 1.0f) {
                    ip[ni] = at(b, 7, t);
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
                ip[ni] = at(a, b, t);
                face[f][nf[f]++] = ip[ni];
                ni++;
            }
        } else {
            ip[ni] = at(0, a, t);
            ni++;
        }
        t = T(a, c);
        if (t >= 0.0f && t <= 1.0f) {
            ip[ni] = at(a, c, t);
            face[f][nf[f]++] = ip[ni];
            ni++;
        }
    }
    // volumeOfPolyhedron: pyramids from v0 over the cut polygon and the three
    // faces (the other faces hold v0 and add nothing).
    float vol = 0.0f;
    vol += turned_pyramid(ip, ni, n, v[0]);
    for (int f = 0; f < 3; f++)
        if (nf[f] > 2) {
            const F3 rn = normalize3(cross(sub(face[f][0], face[f][2]), sub(face[f][1], face[f][0])));
            vol += turned_pyramid(face[f], nf[f], rn, v[0]);
        }
    return vol;
}

// The ratio processVoxel<PartialNode> stores for triangle t (class bb) at the
// voxel it is handed: the padded node coordinates (x + 1, y + adj.y,
// z + adj.z) of process{1D,2D,3D}Triangles (3374-3383, 3489-3495), so the
// box is centred at m + d * (node coordinates).  1D triangles pass a zero
// normal (3389-3407).  NaN results are stored as the quiet NaN 0x7FC00000.
float partial_ratio(const float *V, const u32 *I, u32 t, u32 bb, const int32_t node[3], F3 m,
                    float d) {
    const F3 tri[3] = {fetch_f(V, I, t, 0), fetch_f(V, I, t, 1), fetch_f(V, I, t, 2)};
    const F3 n = bb == BB_1D ? F3{0.0f, 0.0f, 0.0f}
                             : cross(sub(tri[0], tri[2]), sub(tri[1], tri[0]));
    const float vol = cut_volume(tri, node, n, m, d);
    float r = std::fabs(1.0f - vol / (d * d * d));
    if (r != r) {
        const u32 q = 0x7FC00000u;
        std::memcpy(&r, &q, 4);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Nodes.
// ---------------------------------------------------------------------------
// r is PartialNode's ratio: float(solid) from constructNodeList2
// (src/voxelizer.cu:1675), the cut ratio where a triangle wrote the node.
struct NodeRec { uint16_t bid; uint8_t mat; float r; };

// fillNodeList2 repeated until the error flag stays false, per alloc subspace
// (src/voxelizer.cu:1791-1913, src/host_side.cpp:1183-1206).  In place, in
// index order: the literal single-thread restatement.
void erode_inplace(std::vector<NodeRec> &nodes, u32 X, u32 Y, u32 Z, Span ys, Span zs) {
    const size_t npx = X, npy = (size_t)X * Y;
    bool error = true;
    while (error) {
        error = false;
        for (u32 z = zs.min; z < zs.max; z++)
            for (u32 y = ys.min; y < ys.max; y++)
                for (u32 x = 0; x < X; x++) {
                    size_t n = npy * z + npx * y + x;
                    NodeRec &nd = nodes[n];
                    if (nd.bid == 0) continue;
                    uint8_t perm = 0;
                    if (x > 0 && nodes[n - 1].bid > 0) perm += 1;
                    if (x < X - 1 && nodes[n + 1].bid > 0) perm += 2;
                    if (y > 0 && nodes[n - npx].bid > 0) perm += 4;
                    if (y < Y - 1 && nodes[n + npx].bid > 0) perm += 8;
                    if (z > 0 && nodes[n - npy].bid > 0) perm += 16;
                    if (z < Z - 1 && nodes[n + npy].bid > 0) perm += 32;
                    uint8_t nb = kOrient.lut[perm];
                    nd.bid = nb;
                    if (nb == 0) error = true;
                }
    }
}

// Same fixed point computed with snapshot (Jacobi) passes in parallel: the
// removal rule is monotone, so both orders reach the greatest fixed point of
// the subspace with the outside frozen (tests check the two agree).
void erode_parallel(std::vector<NodeRec> &nodes, u32 X, u32 Y, u32 Z, Span ys, Span zs) {
    const size_t npx = X, npy = (size_t)X * Y;
    std::vector<uint8_t> solid(nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) solid[i] = nodes[i].bid > 0;
    bool changed = true;
    std::vector<uint8_t> next;
    while (changed) {
        next = solid;
        long cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
        for (long zz = zs.min; zz < (long)zs.max; zz++) {
            u32 z = (u32)zz;
            for (u32 y = ys.min; y < ys.max; y++)
                for (u32 x = 0; x < X; x++) {
                    size_t n = npy * z + npx * y + x;
                    if (!solid[n]) continue;
                    bool lx = x > 0 && solid[n - 1], rx = x < X - 1 && solid[n + 1];
                    bool iy = y > 0 && solid[n - npx], oy = y < Y - 1 && solid[n + npx];
                    bool dz = z > 0 && solid[n - npy], uz = z < Z - 1 && solid[n + npy];
                    if (!((lx || rx) && (iy || oy) && (dz || uz))) {
                        next[n] = 0;
                        cnt++;
                    }
                }
        }
        changed = cnt > 0;
        solid.swap(next);
    }
#pragma omp parallel for schedule(static)
    for (long zz = zs.min; zz < (long)zs.max; zz++) {
        u32 z = (u32)zz;
        for (u32 y = ys.min; y < ys.max; y++)
            for (u32 x = 0; x < X; x++) {
                size_t n = npy * z + npx * y + x;
                if (nodes[n].bid == 0) continue;
                if (!solid[n]) {
                    nodes[n].bid = 0;
                    continue;
                }
                uint8_t perm = 0;
                if (x > 0 && solid[n - 1]) perm += 1;
                if (x < X - 1 && solid[n + 1]) perm += 2;
                if (y > 0 && solid[n - npx]) perm += 4;
                if (y < Y - 1 && solid[n + npx]) perm += 8;
                if (z > 0 && solid[n - npy]) perm += 16;
                if (z < Z - 1 && solid[n + npy]) perm += 32;
                nodes[n].bid = kOrient.lut[perm];
            }
    }
}

// calculateFCCBoundaries (src/voxelizer.cu:1930-2048), in place in index
// order, one pass per alloc subspace.
void fcc_boundaries(std::vector<NodeRec> &nodes, u32 X, u32 Y, u32 Z, Span ys, Span zs) {
    const size_t npx = X, npy = (size_t)X * Y;
    for (u32 z = zs.min; z < zs.max; z++)
        for (u32 y = ys.min; y < ys.max; y++)
            for (u32 x = 0; x < X; x++) {
                size_t n = npy * z + npx * y + x;
                if (nodes[n].bid == 0) continue;
                uint16_t nb = 0;
                const bool left = x > 0, right = x < X - 1, in = y > 0, out = y < Y - 1;
                const bool down = z > 0, up = z < Z - 1;
                const bool e = (x + z) % 2 == 0;
                const size_t ye = e ? npx : 0, yo = e ? 0 : npx;
                auto s = [&](size_t i) { return nodes[i].bid > 0; };
                if (down && ((e && in) || !e) && s(n - npy - ye)) nb += 1;
                if (down && left && s(n - npy - 1)) nb += 2;
                if (down && right && s(n - npy + 1)) nb += 4;
                if (down && ((!e && out) || e) && s(n - npy + yo)) nb += 8;
                if (left && ((e && in) || !e) && s(n - 1 - ye)) nb += 16;
                if (right && ((e && in) || !e) && s(n + 1 - ye)) nb += 32;
                if (left && ((!e && out) || e) && s(n - 1 + yo)) nb += 64;
                if (right && ((!e && out) || e) && s(n + 1 + yo)) nb += 128;
                if (up && ((e && in) || !e) && s(n + npy - ye)) nb += 256;
                if (up && left && s(n + npy - 1)) nb += 512;
                if (up && right && s(n + npy + 1)) nb += 1024;
                if (up && ((!e && out) || e) && s(n + npy + yo)) nb += 2048;
                nodes[n].bid = nb;
            }
}

// Node packing (include/node_types.h): ShortNode bim = mat<<5 | bid&31 in a
// uchar; LongNode {bid, mat}; ShortFCCNode bim = mat<<12 | bid&4095 in a
// ushort; LongFCCNode {ushort bid, uchar mat, pad}.
void pack_nodes(const std::vector<NodeRec> &nodes, int kind, void *out) {
    size_t n = nodes.size();
    if (kind == VOX_NODE_SHORT) {
        uint8_t *o = (uint8_t *)out;
        for (size_t i = 0; i < n; i++)
            o[i] = (uint8_t)((nodes[i].mat << 5) | (nodes[i].bid & 31));
    } else if (kind == VOX_NODE_LONG) {
        uint8_t *o = (uint8_t *)out;
        for (size_t i = 0; i < n; i++) {
            o[2 * i] = (uint8_t)nodes[i].bid;
            o[2 * i + 1] = nodes[i].mat;
        }
    } else if (kind == VOX_NODE_PARTIAL) {
        // {uchar bid, uchar mat, 2 padding bytes (written 0), float ratio}
        uint8_t *o = (uint8_t *)out;
        for (size_t i = 0; i < n; i++) {
            o[8 * i] = (uint8_t)nodes[i].bid;
            o[8 * i + 1] = nodes[i].mat;
            o[8 * i + 2] = o[8 * i + 3] = 0;
            std::memcpy(o + 8 * i + 4, &nodes[i].r, 4);
        }
    } else if (kind == VOX_NODE_SHORT_FCC) {
        uint16_t *o = (uint16_t *)out;
        for (size_t i = 0; i < n; i++)
            o[i] = (uint16_t)((nodes[i].mat << 12) | (nodes[i].bid & 4095));
    } else {
        uint8_t *o = (uint8_t *)out;
        for (size_t i = 0; i < n; i++) {
            uint16_t b = nodes[i].bid;
            std::memcpy(o + 4 * i, &b, 2);
            o[4 * i + 2] = nodes[i].mat;
            o[4 * i + 3] = 0;
        }
    }
}

// One device's partition of the grid (voxelizeWorker, src/host_side.cpp:
// 901-1244): the partition [pmin, pmax) in (y, z) ext coordinates, its
// neighbour flags (908-950), the ext partition that reaches one row into each
// neighbour (994-1011) and the alloc layout (partition + 2 rows / planes).
PartLayout part_layout(const vox_grid &g, const u32 pmin[2], const u32 pmax[2]) {
    const u32 res[2] = {g.res[1], g.res[2]};
    bool lo[2] = {false, false}, hi[2] = {false, false};  // (left, down), (right, up)
    for (int a = 0; a < 2; a++) {
        if (pmin[a] == 0 && pmax[a] != res[a]) hi[a] = true;
        else if (pmax[a] == res[a] && pmin[a] != 0) lo[a] = true;
        else if (pmin[a] != 0 && pmax[a] != res[a]) lo[a] = hi[a] = true;
    }
    PartLayout L;
    const u32 emin[2] = {pmin[0] - (lo[0] ? 1u : 0u), pmin[1] - (lo[1] ? 1u : 0u)};
    const u32 emax[2] = {pmax[0] + (hi[0] ? 1u : 0u), pmax[1] + (hi[1] ? 1u : 0u)};
    L.Yext = emax[0] - emin[0];
    L.Zext = emax[1] - emin[1];
    L.Yalloc = pmax[0] - pmin[0] + 2;
    L.Zalloc = pmax[1] - pmin[1] + 2;
    L.oy = lo[0] ? 0u : 1u;
    L.oz = lo[1] ? 0u : 1u;
    L.extMin[0] = g.min_vertex[0] + double(0u) * g.voxel_length;
    L.extMin[1] = g.min_vertex[1] + double(emin[0]) * g.voxel_length;
    L.extMin[2] = g.min_vertex[2] + double(emin[1]) * g.voxel_length;
    L.any_neighbour = lo[0] || hi[0] || lo[1] || hi[1];
    return L;
}

// zeroPadding (src/voxelizer.cu:4811-4844), Cartesian nodes.
void zero_border(std::vector<NodeRec> &nodes, u32 X, u32 Y, u32 Z) {
#pragma omp parallel for schedule(static)
    for (long zz = 0; zz < (long)Z; zz++) {
        const u32 z = (u32)zz;
        for (u32 y = 0; y < Y; y++)
            for (u32 x = 0; x < X; x++)
                if (x == 0 || x == X - 1 || y == 0 || y == Y - 1 || z == 0 || z == Z - 1)
                    nodes[((size_t)z * Y + y) * X + x] = {0, 0, 0.0f};
    }
}

// Nodes of one Cartesian partition before classification: the solid bits and
// the surface winners of its ext space, written with its alloc offsets.
// PartialNode (partial): r = float(solid) (constructNodeList2,
// src/voxelizer.cu:1675), and where a triangle wrote the node, its cut ratio
// at the node's alloc coordinates.
std::vector<NodeRec> part_nodes(const float *V, const u32 *I, u32 nt, const uint8_t *mats,
                                const vox_grid &g, const PartLayout &L, const u32 voxSplit[3],
                                const u32 matSplit[3], bool partial) {
    const u32 X = g.alloc[0], Y = L.Yalloc, Z = L.Zalloc;
    const size_t words = (size_t)X / 32 * Y * Z, N = (size_t)X * Y * Z;
    std::vector<u32> bits(words);
    solid_bits_layout(V, I, nt, g, L, voxSplit[1], voxSplit[2], bits.data());
    std::vector<NodeRec> nodes(N);
#pragma omp parallel for schedule(static)
    for (long w = 0; w < (long)words; w++)
        for (u32 b = 0; b < 32; b++) {
            const uint16_t bit = (uint16_t)((bits[w] >> (31 - b)) & 1u);
            nodes[(size_t)w * 32 + b] = {bit, 0, float(bit)};
        }
    if (mats) {
        std::vector<u32> keys(N, 0u);
        SurfTarget st{X, Y, Z, 0, keys.data(), N};
        st.dy = (int32_t)L.oy - 1;
        st.dz = (int32_t)L.oz - 1;
        const u32 extRes[3] = {g.res[0], L.Yext, L.Zext};
        surface_grid_res(V, I, nt, g, L.extMin, extRes, matSplit, st);
        const F3 m = {(float)L.extMin[0], (float)L.extMin[1], (float)L.extMin[2]};
        const float d = (float)g.voxel_length;
        // processVoxel stores Node(1, materials[tri]) (src/voxelizer.cu:2988),
        // PartialNode(1, materials[tri], ratio) (3032).
#pragma omp parallel for schedule(dynamic, 4096)
        for (long i = 0; i < (long)N; i++) {
            if (!keys[i]) continue;
            const u32 t = keys[i] & 0x3FFFFFFFu;
            float r = nodes[i].r;
            if (partial) {
                const int32_t node[3] = {(int32_t)(i % X), (int32_t)((i / X) % Y),
                                         (int32_t)(i / ((size_t)X * Y))};
                r = partial_ratio(V, I, t, keys[i] >> 30, node, m, d);
            }
            nodes[i] = {1, mats[t], r};
        }
    }
    return nodes;
}

// Erosion per alloc subspace, then the padding cleared when the partition
// has neighbours (src/host_side.cpp:1183-1220).
void classify_part(std::vector<NodeRec> &nodes, u32 X, u32 Y, u32 Z, const u32 voxSplit[3],
                   bool any_neighbour, bool literal) {
    std::vector<Span> ys = split_axis_max(Y, voxSplit[1]), zs = split_axis_max(Z, voxSplit[2]);
    for (const Span &z : zs)
        for (const Span &y : ys) {
            if (literal) erode_inplace(nodes, X, Y, Z, y, z);
            else erode_parallel(nodes, X, Y, Z, y, z);
        }
    if (any_neighbour) zero_border(nodes, X, Y, Z);
}

// rotateVertices (src/host_side.cpp:3261-3290): (x, y) -> (a - y, b + x) with
// a = c.y + c.x, b = c.y - c.x and c = (max - min) / 2 of the bounding box
// (calculateBoundingBox's strict comparisons, 2462-2488).
std::vector<float> rotate_vertices(const float *V, u32 nv) {
    std::vector<float> r(V, V + 3 * (size_t)nv);
    if (nv == 0) return r;
    double mn[3] = {V[0], V[1], V[2]}, mx[3] = {V[0], V[1], V[2]};
    for (u32 i = 1; i < nv; i++)
        for (int c = 0; c < 3; c++) {
            if (V[3 * i + c] > mx[c]) mx[c] = V[3 * i + c];
            if (V[3 * i + c] < mn[c]) mn[c] = V[3 * i + c];
        }
    const double cx = (mx[0] - mn[0]) / 2.0, cy = (mx[1] - mn[1]) / 2.0;
    const double a = cy + cx, b = cy - cx;
    for (u32 i = 0; i < nv; i++) {
        const double tx = r[3 * i];
        r[3 * i] = float(a - double(r[3 * i + 1]));
        r[3 * i + 1] = float(b + tx);
    }
    return r;
}

// unRotateNodes (src/voxelizer.cu:2061-2124), Cartesian: node (x, y, z) of a
// (X, Y, Z) array goes to (y, X - 1 - x, z) of a (Y, X, Z) array.
std::vector<NodeRec> unrotate_nodes(const std::vector<NodeRec> &in, u32 X, u32 Y, u32 Z) {
    std::vector<NodeRec> out(in.size());
    for (u32 z = 0; z < Z; z++)
        for (u32 y = 0; y < Y; y++)
            for (u32 x = 0; x < X; x++)
                out[(size_t)X * Y * z + (size_t)Y * (X - 1 - x) + y] = in[((size_t)z * Y + y) * X + x];
    return out;
}

int voxelize_nodes(const float *V, u32 nv, const u32 *I, u32 nt, const uint8_t *mats,
                   const vox_grid &g, int kind, const u32 voxSplit[3], const u32 matSplit[3],
                   void *out, bool literal) {
    (void)nv;
    const bool fcc = kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
    if (kind != VOX_NODE_SHORT && kind != VOX_NODE_LONG && kind != VOX_NODE_PARTIAL && !fcc)
        return 1;
    const u32 X = g.alloc[0], Y = g.alloc[1], Z = g.alloc[2];
    const size_t words = (size_t)X / 32 * Y * Z;
    std::vector<u32> bits(words);
    const double base[3] = {g.min_vertex[0], g.min_vertex[1], g.min_vertex[2]};
    if (!fcc) {
        // voxelizeWorker (src/host_side.cpp:901-1244) on the whole grid.
        const PartLayout L = whole_layout(g, base);
        std::vector<NodeRec> nodes =
            part_nodes(V, I, nt, mats, g, L, voxSplit, matSplit, kind == VOX_NODE_PARTIAL);
        classify_part(nodes, X, Y, Z, voxSplit, false, literal);
        pack_nodes(nodes, kind, out);
        return 0;
    }
    // fccWorker (src/host_side.cpp:1262-1698).
    const u32 FX = 2 * X, FY = Y, FZ = 2 * Z;
    const size_t N = (size_t)FX * FY * FZ;
    std::vector<NodeRec> nodes(N);
    const double h = g.voxel_length / 2.0;
    const double shifts[4][3] = {{0.0, 0.0, 0.0}, {h, h, 0.0}, {0.0, h, h}, {h, 0.0, h}};
    for (int gt = 1; gt <= 4; gt++) {
        double em[3] = {base[0] + shifts[gt - 1][0], base[1] + shifts[gt - 1][1],
                        base[2] + shifts[gt - 1][2]};
        solid_bits(V, I, nt, g, em, voxSplit[1], voxSplit[2], bits.data());
        // convertToFCCGrid (src/voxelizer.cu:1711-1776).
        const u32 ox = (gt == 2 || gt == 4), oz = (gt == 3 || gt == 4);
#pragma omp parallel for schedule(static)
        for (long zz = 0; zz < (long)Z; zz++) {
            u32 z = (u32)zz;
            for (u32 y = 0; y < Y; y++)
                for (u32 x = 0; x < X; x++) {
                    size_t w = ((size_t)z * Y + y) * (X / 32) + x / 32;
                    uint16_t bit = (uint16_t)((bits[w] >> (31 - x % 32)) & 1u);
                    size_t idx = (size_t)FX * FY * (2 * z + oz) + (size_t)FX * y + 2 * x + ox;
                    nodes[idx] = {bit, 0, 0.0f};
                }
        }
    }
    if (mats) {
        std::vector<u32> keys(N, 0u);
        for (int gt = 1; gt <= 4; gt++) {
            double em[3] = {base[0] + shifts[gt - 1][0], base[1] + shifts[gt - 1][1],
                            base[2] + shifts[gt - 1][2]};
            SurfTarget st{X, Y, Z, gt, keys.data(), N};
            surface_grid(V, I, nt, g, em, matSplit, st);
        }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)N; i++)
            if (keys[i]) nodes[i] = {1, mats[keys[i] & 0x3FFFFFFFu], 0.0f};
    }
    // allocYzSplits computed on the per-grid alloc partition, then z doubled
    // (src/host_side.cpp:1415-1419, 1500-1511).
    std::vector<Span> ys = split_axis_max(Y, voxSplit[1]), zs = split_axis_max(Z, voxSplit[2]);
    for (const Span &z : zs)
        for (const Span &y : ys) fcc_boundaries(nodes, FX, FY, FZ, y, {2 * z.min, 2 * z.max});
    pack_nodes(nodes, kind, out);
    return 0;
}


// ---------------------------------------------------------------------------
// Fragile decisions: where the original CUDA build could legitimately end
// with different nodes than this restatement.
//  * nvcc contracts a*b + c into one FMA by default, so a test value within a
//    few ulps of its threshold can fall on the other side there.  A decision
//    is listed when its value lies within eps times the magnitude of its
//    terms (eps64 for the FP64 solid tests, eps32 for the FP32 surface
//    tests), and so are the voxels it decides.
//  * processVoxel's plain store races (src/voxelizer.cu:2988): a voxel
//    written by several triangles of the highest class (kernels run 1D, 2D,
//    3D in order) with different materials may keep any of them.
// Per-voxel flags: 1 solid state rests on a fragile solid decision, 2 surface
// membership rests on a fragile surface decision, 4 written by several
// triangles, 8 several materials among the highest-class writers.
// ---------------------------------------------------------------------------
enum : uint8_t { FR_SOLID = 1, FR_SURF = 2, FR_MULTI = 4, FR_RACE = 8 };

// Values that are exactly on their threshold are not listed: they come from
// structured geometry (planes on voxel boundaries, centres on edges, the
// cases the DBL_MIN tie break settles) whose products are exact, and a
// contracted evaluation computes the same exact value.
inline bool near_zero(double v, double scale, double eps) {
    return v != 0.0 && std::fabs(v) <= eps * scale;
}
// x within eps of an integer (relative to |x| + 1), but not on it.
inline bool near_int(double x, double eps) {
    const double r = std::fabs(x - std::nearbyint(x));
    return r != 0.0 && r <= eps * (std::fabs(x) + 1.0);
}

struct FragileMap {
    uint8_t *flags;
    size_t n;
    // node index of padded voxel (x, y, z) of a Cartesian grid or FCC sub-grid
    u32 X, Y, Z;
    int gridType;
    size_t index(u32 x, u32 y, u32 z) const {
        if (gridType == 0) return ((size_t)z * Y + y) * X + x;
        size_t l = 2 * (size_t)X, a = l * Y;
        size_t ox = gridType == 2 || gridType == 4, oz = gridType == 3 || gridType == 4;
        return a * (2 * (size_t)z + oz) + l * y + 2 * (size_t)x + ox;
    }
    void mark(u32 x, u32 y, u32 z, uint8_t f) const {
        size_t i = index(x, y, z);
        if (i < n) __atomic_fetch_or(&flags[i], f, __ATOMIC_RELAXED);
    }
};

// Solid decisions of one subspace (perform_voxelization's queue and fill).
// Counts: [0] fill tests evaluated, [1] of them fragile, [2] fragile crossing
// positions.
void fragile_solid_subspace(const SolidCtx &c, Span ys, Span zs, double eps,
                            const FragileMap &fm, uint64_t counts[3]) {
    const u32 smin[2] = {ys.min, zs.min}, smax[2] = {ys.max, zs.max};
    uint64_t n_tests = 0, n_frag = 0, n_cross = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n_tests, n_frag, n_cross)
    for (long tl = 0; tl < (long)c.nt; tl++) {
        const u32 t = (u32)tl;
        D3 v[3] = {fetch_d(c.V, c.I, t, 0), fetch_d(c.V, c.I, t, 1), fetch_d(c.V, c.I, t, 2)};
        TileSetup ts;
        if (!tile_setup(v, c.extMin, c.d, smin, smax, ts)) continue;
        const double tN = (v[0].y * v[1].z + v[2].y * v[0].z + v[2].z * v[1].y) -
                          (v[2].y * v[1].z + v[0].z * v[1].y + v[2].z * v[0].y);
        const double tNs = std::fabs(v[0].y * v[1].z) + std::fabs(v[2].y * v[0].z) +
                           std::fabs(v[2].z * v[1].y) + std::fabs(v[2].y * v[1].z) +
                           std::fabs(v[0].z * v[1].y) + std::fabs(v[2].z * v[0].y);
        const bool tri_fragile = near_zero(tN, tNs, eps);
        const FillSetup fs = fill_setup(v);
        for (u32 ty = ts.tmin[0]; ty <= ts.tmax[0]; ty++)
            for (u32 tz = ts.tmin[1]; tz <= ts.tmax[1]; tz++) {
                // Filter: queued, and queued only through fragile centres?
                bool hit = false, solid_hit = false;
                for (u32 i = ty * 4; i < ty * 4 + 4 && i < ts.cmax[0]; i++)
                    for (u32 j = tz * 4; j < tz * 4 + 4 && j < ts.cmax[1]; j++) {
                        if (i < ts.cmin[0] || j < ts.cmin[1]) continue;
                        D2 p = {double(i) * c.d + c.extMin[1], double(j) * c.d + c.extMin[2]};
                        bool r = true, sure = true;
                        for (int k = 0; k < 3; k++) {
                            double a = ts.ne[k].x * p.x, b = ts.ne[k].y * p.y;
                            double val = a + b + ts.de[k];
                            r &= val >= 0.0;
                            sure &= !near_zero(val, std::fabs(a) + std::fabs(b) + std::fabs(ts.de[k]), eps);
                        }
                        if (r) {
                            hit = true;
                            solid_hit |= sure;
                        }
                    }
                if (!hit) continue;
                const bool queue_fragile = tri_fragile || !solid_hit;
                for (int lane = 0; lane < 16; lane++) {
                    const u32 cy = u32(lane % 4) + ty * 4, cz = u32(lane / 4) + tz * 4;
                    if (cy < smin[0] || cy >= smax[0] || cz < smin[1] || cz >= smax[1]) continue;
                    D2 p = {double(cy) * c.d + c.extMin[1], double(cz) * c.d + c.extMin[2]};
                    u32 ok = 0;
                    bool frag = queue_fragile;
                    for (int i = 0; i < 3; i++) {
                        double a = fs.ne[i].x * p.x, b = fs.ne[i].y * p.y;
                        double r = a + b + fs.de[i];
                        frag |= near_zero(r, std::fabs(a) + std::fabs(b) + std::fabs(fs.de[i]), eps);
                        if (fs.ne[i].x > 0.0 || (fs.ne[i].x == 0.0 && fs.ne[i].y < 0.0)) r += DBL_MIN;
                        if (r > 0.0) ok++;
                    }
                    n_tests++;
                    if (ok != 3 && !frag) continue;
                    // Where the crossing is (or would be).
                    D3 A = {fs.v0.x - c.extMin[0], fs.v0.y - p.x, fs.v0.z - p.y};
                    double B = A.x * fs.n.x + A.y * fs.n.y + A.z * fs.n.z;
                    double px = B / fs.n.x;
                    if (!(px == px)) continue;
                    px = px < 0.0 ? 0.0 : px;
                    const double q = px / c.d;
                    const u32 X0 = sat_u32(std::ceil(q)) + 1u;
                    if (frag) {
                        n_frag++;
                        for (u32 x = X0; x < c.X; x++) fm.mark(x, cy + 1, cz + 1, FR_SOLID);
                    } else if (near_int(q, eps)) {
                        // ceil(q) may land one voxel lower or higher: the voxel
                        // between the two crossing positions changes state.
                        n_cross++;
                        const double k = std::nearbyint(q);
                        const u32 xv = sat_u32(k) + 1u;
                        if (xv < c.X) fm.mark(xv, cy + 1, cz + 1, FR_SOLID);
                    }
                }
            }
    }
    counts[0] += n_tests;
    counts[1] += n_frag;
    counts[2] += n_cross;
}

// Surface decisions of one triangle in one subspace, mirroring
// surface_triangle.  Counts: [0] tests evaluated, [1] fragile.  Every write
// also goes to the writer record of the voxel (rec: class << 10 | count << 8
// | material, count saturating at 3; a 0x80000000 bit marks two materials
// among the highest-class writers).
struct WriterRec {
    u32 *rec;
    size_t n;
};
inline void record_writer(const WriterRec &w, size_t i, u32 cls, u32 mat) {
    if (i >= w.n) return;
    u32 cur = __atomic_load_n(&w.rec[i], __ATOMIC_RELAXED);
    while (true) {
        u32 nxt;
        const u32 cnt = (cur >> 8) & 3u;
        if (cnt == 0) {
            nxt = (cls << 10) | (1u << 8) | mat;
        } else {
            const u32 ccls = (cur >> 10) & 3u, cmat = cur & 0xFFu;
            const u32 ncnt = cnt < 3 ? cnt + 1 : 3;
            if (cls > ccls) nxt = (cls << 10) | (ncnt << 8) | mat;
            else if (cls == ccls)
                nxt = (cur & ~(3u << 8)) | (ncnt << 8) | (mat != cmat ? 0x80000000u : 0u);
            else nxt = (cur & ~(3u << 8)) | (ncnt << 8);
        }
        if (__atomic_compare_exchange_n(&w.rec[i], &cur, nxt, true, __ATOMIC_RELAXED,
                                        __ATOMIC_RELAXED))
            return;
    }
}

// A point's projection on plane P: 0 = (y, z), 1 = (z, x), 2 = (x, y).
inline F2 proj(const float p[3], int P) {
    return P == 0 ? F2{p[1], p[2]} : P == 1 ? F2{p[2], p[0]} : F2{p[0], p[1]};
}
// Whether one of the three edge values of a projected test lies within eps
// of 0 relative to its terms.
inline bool frag_test(const OverlapData &o, F2 p2, double eps) {
    bool f = false;
    for (int i = 0; i < 3; i++) {
        const float a = o.ne[i].x * p2.x, b = o.ne[i].y * p2.y;
        const float v = (a + b) + o.de[i];
        f |= near_zero((double)v, std::fabs((double)a) + std::fabs((double)b) +
                                      std::fabs((double)o.de[i]), eps);
    }
    return f;
}

void fragile_surface_triangle(const F3 t[3], u32 type, u32 mat, F3 m, float d, const u32 smin[3],
                              const u32 smax[3], const SurfTarget &st, const FragileMap &fm,
                              const WriterRec &wr, double eps, uint64_t &n_tests,
                              uint64_t &n_frag) {
    const int bb = (int)(type >> 30);
    const int axis = (int)((type >> 28) & 3u);
    F3 tmn, tmx;
    tri_bounds(t, tmn, tmx);
    VoxBB vb = voxel_bounds_half(tmn, tmx, m, d);
    const float mm[3] = {m.x, m.y, m.z};
    const float lo3[3] = {(tmn.x - m.x) / d - 0.5f, (tmn.y - m.y) / d - 0.5f, (tmn.z - m.z) / d - 0.5f};
    const float hi3[3] = {(tmx.x - m.x) / d + 0.5f, (tmx.y - m.y) / d + 0.5f, (tmx.z - m.z) / d + 0.5f};
    // The class, axis and voxel range rest on six roundings and, for 3D
    // triangles, on the dominant-axis comparisons.
    bool tri_frag = false;
    for (int a = 0; a < 3; a++) tri_frag |= near_int(lo3[a], eps) || near_int(hi3[a], eps);
    const F3 n = cross(sub(t[0], t[2]), sub(t[1], t[0]));
    const float nn[3] = {n.x, n.y, n.z};
    if (bb == BB_3D) {
        const float ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
        const double big = std::max(ax, std::max(ay, az));
        tri_frag |= near_zero((double)ax - ay, big, eps) || near_zero((double)ax - az, big, eps) ||
                    near_zero((double)ay - az, big, eps);
    }
    if (!adjust_range(vb, smin, smax)) return;
    const int32_t vlo[3] = {vb.min.x, vb.min.y, vb.min.z}, vhi[3] = {vb.max.x, vb.max.y, vb.max.z};
    auto visit = [&](const int32_t idx[3], bool written, bool frag) {
        const int32_t x = idx[0] + 1, y = idx[1] + 1, z = idx[2] + 1;
        if (written) {
            const size_t i = st.index(x, y, z);
            if (i < st.nkeys) record_writer(wr, i, (u32)bb, mat);
        }
        if ((written || frag) && (frag || tri_frag)) fm.mark((u32)x, (u32)y, (u32)z, FR_SURF);
    };
    if (bb == BB_1D) {
        int32_t idx[3] = {vlo[0], vlo[1], vlo[2]};
        for (idx[axis] = vlo[axis]; idx[axis] <= vhi[axis]; idx[axis]++) visit(idx, true, false);
        return;
    }
    const OverlapData ov[3] = {setup_yz(t, n, d), setup_zx(t, n, d), setup_xy(t, n, d)};
    const int pd = axis == AX_X ? 0 : axis == AX_Y ? 1 : 2;  // plane normal to the axis
    const int u = axis == AX_X ? AX_Y : AX_X, w = axis == AX_Z ? AX_Y : AX_Z;
    int32_t idx[3];
    float p[3];
    for (idx[u] = vlo[u]; idx[u] <= vhi[u]; idx[u]++)
        for (idx[w] = vlo[w]; idx[w] <= vhi[w]; idx[w]++) {
            p[u] = vox_corner(idx[u], mm[u], d);
            p[w] = vox_corner(idx[w], mm[w], d);
            p[axis] = 0.0f;
            n_tests++;
            const bool fcol = frag_test(ov[pd], proj(p, pd), eps);
            n_frag += fcol;
            const bool in = test2(ov[pd], proj(p, pd));
            if (bb == BB_2D) {
                idx[axis] = vlo[axis];
                visit(idx, in, fcol);
                continue;
            }
            if (!in && !fcol) continue;
            // Depth range of the column (determineDepthRange*), with the
            // margins of its two end roundings.
            float pmin[3] = {p[0], p[1], p[2]}, pmax[3] = {p[0], p[1], p[2]};
            for (int c = 0; c < 3; c++) {
                if (c == axis) continue;
                const bool pos = nn[c] > 0.0f;
                pmin[c] = p[c] + float(pos) * d;
                pmax[c] = p[c] + float(!pos) * d;
            }
            auto depth = [&](const float q[3], double &scale) {
                const F3 e = {t[0].x - q[0], t[0].y - q[1], t[0].z - q[2]};
                scale = (std::fabs((double)e.x * n.x) + std::fabs((double)e.y * n.y) +
                         std::fabs((double)e.z * n.z)) / std::fabs((double)nn[axis]);
                return dot(e, n) / nn[axis];
            };
            double sa, sb;
            float a = depth(pmin, sa), b = depth(pmax, sb);
            if (a > b) {
                std::swap(a, b);
                std::swap(sa, sb);
            }
            const float qa = (a - mm[axis]) / d - 0.5f, qb = (b - mm[axis]) / d + 0.5f;
            const int32_t lo0 = vox_cbound_min(a, mm[axis], d), hi0 = vox_cbound_max(b, mm[axis], d);
            const double ra = std::fabs(qa - std::nearbyint(qa)), rb = std::fabs(qb - std::nearbyint(qb));
            const bool fa = ra != 0.0 && ra <= eps * (std::fabs(qa) + 1.0 + sa / d);
            const bool fb = rb != 0.0 && rb <= eps * (std::fabs(qb) + 1.0 + sb / d);
            n_tests++;
            n_frag += fa || fb;
            const int32_t lo = std::max(lo0 - (fa ? 1 : 0), vlo[axis]);
            const int32_t hi = std::min(hi0 + (fb ? 1 : 0), vhi[axis]);
            const int32_t rlo = std::max(lo0, vlo[axis]), rhi = std::min(hi0, vhi[axis]);
            for (idx[axis] = lo; idx[axis] <= hi; idx[axis]++) {
                const bool inrange = idx[axis] >= rlo && idx[axis] <= rhi;
                const bool endf = (fa && (idx[axis] == lo0 || idx[axis] == lo0 - 1)) ||
                                  (fb && (idx[axis] == hi0 || idx[axis] == hi0 + 1));
                p[axis] = vox_corner(idx[axis], mm[axis], d);
                bool write = inrange && rlo == rhi;
                bool f = fcol || endf;
                if (inrange && !write) {
                    n_tests++;
                    const int p1 = (pd + 1) % 3, p2 = (pd + 2) % 3;
                    const bool fi = frag_test(ov[p1], proj(p, p1), eps) ||
                                    frag_test(ov[p2], proj(p, p2), eps);
                    n_frag += fi;
                    f |= fi;
                    write = test2(ov[p1], proj(p, p1)) && test2(ov[p2], proj(p, p2));
                }
                visit(idx, in && write, f);
            }
        }
}

// Whole analysis for one grid (Cartesian, or FCC sub-grid gridType 1..4).
void fragile_grid(const float *V, const u32 *I, u32 nt, const uint8_t *mats, const vox_grid &g,
                  const double extMin[3], int gridType, const u32 voxSplit[3],
                  const u32 matSplit[3], double eps64, double eps32, uint8_t *flags, size_t N,
                  u32 *writers, uint64_t counts[5]) {
    SolidCtx c;
    c.V = V;
    c.I = I;
    c.nt = nt;
    for (int a = 0; a < 3; a++) c.extMin[a] = extMin[a];
    c.d = g.voxel_length;
    c.X = g.res[0];
    c.Yext = g.res[1];
    c.Zext = g.res[2];
    c.xSplits = g.x_splits;
    c.bits = nullptr;
    c.Yalloc = g.alloc[1];
    c.Zalloc = g.alloc[2];
    FragileMap fm{flags, N, g.alloc[0], g.alloc[1], g.alloc[2], gridType};
    std::vector<Span> ys = split_axis_max(c.Yext, voxSplit[1]), zs = split_axis_max(c.Zext, voxSplit[2]);
    for (const Span &z : zs)
        for (const Span &y : ys) fragile_solid_subspace(c, y, z, eps64, fm, counts);
    if (!mats) return;
    const F3 m = {(float)extMin[0], (float)extMin[1], (float)extMin[2]};
    const float d = (float)g.voxel_length;
    std::vector<u32> types(nt);
    for (u32 t = 0; t < nt; t++) {
        F3 tri[3] = {fetch_f(V, I, t, 0), fetch_f(V, I, t, 1), fetch_f(V, I, t, 2)};
        types[t] = encode_type(classify(tri, m, d));
    }
    SurfTarget st{g.alloc[0], g.alloc[1], g.alloc[2], gridType, nullptr, N};
    WriterRec wr{writers, N};
    std::vector<Span> xs = split_axis_max(g.res[0], matSplit[0]);
    std::vector<Span> sys = split_axis_max(g.res[1], matSplit[1]);
    std::vector<Span> szs = split_axis_max(g.res[2], matSplit[2]);
    uint64_t n_tests = 0, n_frag = 0;
    for (const Span &z : szs)
        for (const Span &y : sys)
            for (const Span &x : xs) {
                const u32 smin[3] = {x.min, y.min, z.min}, smax[3] = {x.max, y.max, z.max};
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : n_tests, n_frag)
                for (long t = 0; t < (long)nt; t++) {
                    if ((types[t] >> 30) == BB_DEGEN) continue;
                    F3 tri[3] = {fetch_f(V, I, (u32)t, 0), fetch_f(V, I, (u32)t, 1),
                                 fetch_f(V, I, (u32)t, 2)};
                    fragile_surface_triangle(tri, types[t], mats[t], m, d, smin, smax, st, fm, wr,
                                             eps32, n_tests, n_frag);
                }
            }
    counts[3] += n_tests;
    counts[4] += n_frag;
}

}  // namespace

extern "C" {

uint8_t orc_orientation(uint8_t permutation) { return kOrient.lut[permutation & 63]; }

int orc_plan_grid(const float *v, uint32_t nv, int mode, double param, uint32_t vsx,
                  vox_grid *g) {
    if (nv == 0 || vsx == 0) return 1;
    // calculateBoundingBox (src/host_side.cpp:2462-2488).
    double mn[3] = {v[0], v[1], v[2]}, mx[3] = {v[0], v[1], v[2]};
    for (uint32_t i = 1; i < nv; i++)
        for (int c = 0; c < 3; c++) {
            if (v[3 * i + c] > mx[c]) mx[c] = v[3 * i + c];
            if (v[3 * i + c] < mn[c]) mn[c] = v[3 * i + c];
        }
    double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    u32 rx, ry, rz;
    double d;
    if (mode == 0) {
        // determineDimensions() (src/host_side.cpp:2497-2580).
        u32 R = (u32)param;
        if (dx > dy) {
            if (dx > dz) {
                d = dx / double(R - 1);
                ry = sat_u32(std::ceil(dy / d)) + 1;
                rz = sat_u32(std::ceil(dz / d)) + 1;
                rx = R + 2;
            } else {
                d = dz / double(R - 1);
                rz = R;
                rx = sat_u32(std::ceil(dx / d)) + 1;
                ry = sat_u32(std::ceil(dy / d)) + 1;
                rx += 2;
            }
        } else {
            if (dy > dz) {
                d = dy / double(R - 1);
                ry = R;
                rx = sat_u32(std::ceil(dx / d)) + 1;
                rz = sat_u32(std::ceil(dz / d)) + 1;
                rx += 2;
            } else {
                d = dz / double(R - 1);
                rz = R;
                rx = sat_u32(std::ceil(dx / d)) + 1;
                ry = sat_u32(std::ceil(dy / d)) + 1;
                rx += 2;
            }
        }
    } else {
        // determineDimensions(double d) (src/host_side.cpp:2592-2609).
        d = param;
        rx = sat_u32(std::ceil(dx / d)) + 1;
        ry = sat_u32(std::ceil(dy / d)) + 1;
        rz = sat_u32(std::ceil(dz / d)) + 1;
        rx += 2;
    }
    // splitResolutionByMaxDim(voxSplitRes.x) + adjustResolution
    // (src/host_side.cpp:730-738, 2620-2634).
    u32 xs = (rx + vsx - 1) / vsx;
    u32 xmod = 32u * xs;
    if (rx % xmod != 0) rx += xmod - rx % xmod;
    for (int c = 0; c < 3; c++) {
        g->min_vertex[c] = mn[c];
        g->max_vertex[c] = mx[c];
    }
    g->voxel_length = d;
    g->res[0] = rx;
    g->res[1] = ry;
    g->res[2] = rz;
    g->alloc[0] = rx;
    g->alloc[1] = ry + 2;
    g->alloc[2] = rz + 2;
    g->x_splits = xs;
    g->pad_ = 0;
    return 0;
}

int orc_solid_bits(const float *vertices, const uint32_t *indices, uint32_t nt,
                   const vox_grid *grid, const double ext_min[3], uint32_t split_y,
                   uint32_t split_z, uint32_t *bits) {
    if (!grid || grid->x_splits == 0 || split_y == 0 || split_z == 0) return 1;
    solid_bits(vertices, indices, nt, *grid, ext_min, split_y, split_z, bits);
    return 0;
}

int orc_voxelize_nodes(const float *vertices, uint32_t nv, const uint32_t *indices,
                       uint32_t nt, const uint8_t *materials, const vox_grid *grid,
                       int node_kind, const uint32_t vox_split[3], const uint32_t mat_split[3],
                       void *nodes, int n_threads) {
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    bool literal = n_threads == 1;
    int r = voxelize_nodes(vertices, nv, indices, nt, materials, *grid, node_kind, vox_split,
                           mat_split, nodes, literal);
#ifdef _OPENMP
    omp_set_num_threads(prev);
#endif
    return r;
}

int orc_surface_keys(const float *vertices, const uint32_t *indices, uint32_t nt,
                     const vox_grid *grid, const double ext_min[3], const uint32_t mat_split[3],
                     uint32_t *keys) {
    const vox_grid &g = *grid;
    size_t N = (size_t)g.alloc[0] * g.alloc[1] * g.alloc[2];
    std::memset(keys, 0, N * 4);
    SurfTarget st{g.alloc[0], g.alloc[1], g.alloc[2], 0, keys, N};
    surface_grid(vertices, indices, nt, g, ext_min, mat_split, st);
    return 0;
}

void orc_classify(const float *vertices, const uint32_t *indices, uint32_t nt,
                  const float model_min[3], float d, uint32_t *out) {
    F3 m = {model_min[0], model_min[1], model_min[2]};
    for (u32 t = 0; t < nt; t++) {
        F3 tri[3] = {fetch_f(vertices, indices, t, 0), fetch_f(vertices, indices, t, 1),
                     fetch_f(vertices, indices, t, 2)};
        out[t] = encode_type(classify(tri, m, d));
    }
}

int orc_center_crossing(const float tri[9], const double ext_min[3], double d, uint32_t cy,
                        uint32_t cz, uint32_t *voxel_x) {
    D3 v[3];
    for (int k = 0; k < 3; k++) v[k] = {tri[3 * k], tri[3 * k + 1], tri[3 * k + 2]};
    FillSetup fs = fill_setup(v);
    u32 vx = 0;
    bool ok = fill_center(fs, ext_min, d, cy, cz, vx);
    if (voxel_x) *voxel_x = vx;
    return ok ? 1 : 0;
}

int orc_erode_nodes(const uint8_t *solid, uint32_t X, uint32_t Y, uint32_t Z, uint32_t split_y,
                    uint32_t split_z, int literal, uint8_t *bid_out) {
    if (!split_y || !split_z) return 1;
    const size_t N = (size_t)X * Y * Z;
    std::vector<NodeRec> nodes(N);
    for (size_t i = 0; i < N; i++) nodes[i] = {(uint16_t)(solid[i] ? 1 : 0), 0, 0.0f};
    std::vector<Span> ys = split_axis_max(Y, split_y), zs = split_axis_max(Z, split_z);
    for (const Span &z : zs)
        for (const Span &y : ys) {
            if (literal) erode_inplace(nodes, X, Y, Z, y, z);
            else erode_parallel(nodes, X, Y, Z, y, z);
        }
    for (size_t i = 0; i < N; i++) bid_out[i] = (uint8_t)nodes[i].bid;
    return 0;
}

int orc_fcc_boundaries(const uint8_t *solid, uint32_t FX, uint32_t FY, uint32_t FZ,
                       uint16_t *bid_out) {
    const size_t N = (size_t)FX * FY * FZ;
    std::vector<NodeRec> nodes(N);
    for (size_t i = 0; i < N; i++) nodes[i] = {(uint16_t)(solid[i] ? 1 : 0), 0, 0.0f};
    fcc_boundaries(nodes, FX, FY, FZ, {0, FY}, {0, FZ});
    for (size_t i = 0; i < N; i++) bid_out[i] = nodes[i].bid;
    return 0;
}

int orc_fragile(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                const uint8_t *materials, const vox_grid *grid, int node_kind,
                const uint32_t vox_split[3], const uint32_t mat_split[3], double eps64,
                double eps32, uint8_t *flags, uint64_t counts[8]) {
    (void)nv;
    const bool fcc = node_kind == VOX_NODE_SHORT_FCC || node_kind == VOX_NODE_LONG_FCC;
    if (!fcc && node_kind != VOX_NODE_SHORT && node_kind != VOX_NODE_LONG) return 1;
    const vox_grid &g = *grid;
    const size_t N = (size_t)g.alloc[0] * g.alloc[1] * g.alloc[2] * (fcc ? 4 : 1);
    std::memset(flags, 0, N);
    for (int k = 0; k < 8; k++) counts[k] = 0;
    std::vector<u32> writers(materials ? N : 0, 0u);
    const double base[3] = {g.min_vertex[0], g.min_vertex[1], g.min_vertex[2]};
    if (!fcc) {
        fragile_grid(vertices, indices, nt, materials, g, base, 0, vox_split, mat_split, eps64,
                     eps32, flags, N, writers.data(), counts);
    } else {
        const double h = g.voxel_length / 2.0;
        const double shifts[4][3] = {{0.0, 0.0, 0.0}, {h, h, 0.0}, {0.0, h, h}, {h, 0.0, h}};
        for (int gt = 1; gt <= 4; gt++) {
            const double em[3] = {base[0] + shifts[gt - 1][0], base[1] + shifts[gt - 1][1],
                                  base[2] + shifts[gt - 1][2]};
            fragile_grid(vertices, indices, nt, materials, g, em, gt, vox_split, mat_split, eps64,
                         eps32, flags, N, writers.data(), counts);
        }
    }
    uint64_t multi = 0, race = 0;
    if (materials) {
#pragma omp parallel for reduction(+ : multi, race) schedule(static)
        for (long i = 0; i < (long)N; i++) {
            const u32 r = writers[i];
            uint8_t f = 0;
            if (((r >> 8) & 3u) >= 2u) {
                f |= FR_MULTI;
                multi++;
            }
            if (r & 0x80000000u) {
                f |= FR_RACE;
                race++;
            }
            flags[i] |= f;
        }
    }
    counts[5] = multi;
    counts[6] = race;
    uint64_t any = 0;
#pragma omp parallel for reduction(+ : any) schedule(static)
    for (long i = 0; i < (long)N; i++) any += flags[i] != 0;
    counts[7] = any;
    return 0;
}

// Plain output of one partition: voxelizeWorker returns after the solid
// stage when no nodes are asked for (src/host_side.cpp:1040-1055).
int orc_solid_bits_part(const float *vertices, const uint32_t *indices, uint32_t nt,
                        const vox_grid *grid, uint32_t split_y, uint32_t split_z,
                        const uint32_t part_min[2], const uint32_t part_max[2], uint32_t *bits) {
    if (!grid || grid->x_splits == 0 || split_y == 0 || split_z == 0) return 1;
    const vox_grid &g = *grid;
    for (int a = 0; a < 2; a++)
        if (part_min[a] >= part_max[a] || part_max[a] > g.res[1 + a]) return 2;
    solid_bits_layout(vertices, indices, nt, g, part_layout(g, part_min, part_max), split_y,
                      split_z, bits);
    return 0;
}

int orc_voxelize_part(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                      const uint8_t *materials, const vox_grid *grid, int node_kind,
                      const uint32_t vox_split[3], const uint32_t mat_split[3],
                      const uint32_t part_min[2], const uint32_t part_max[2], void *nodes_out,
                      int n_threads) {
    (void)nv;
    if (node_kind != VOX_NODE_SHORT && node_kind != VOX_NODE_LONG && node_kind != VOX_NODE_PARTIAL)
        return 1;
    const vox_grid &g = *grid;
    for (int a = 0; a < 2; a++)
        if (part_min[a] >= part_max[a] || part_max[a] > g.res[1 + a]) return 2;
#ifdef _OPENMP
    int prev = omp_get_max_threads();
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    const PartLayout L = part_layout(g, part_min, part_max);
    std::vector<NodeRec> nodes =
        part_nodes(vertices, indices, nt, materials, g, L, vox_split, mat_split,
                   node_kind == VOX_NODE_PARTIAL);
    classify_part(nodes, g.alloc[0], L.Yalloc, L.Zalloc, vox_split, L.any_neighbour,
                  n_threads == 1);
    pack_nodes(nodes, node_kind, nodes_out);
#ifdef _OPENMP
    omp_set_num_threads(prev);
#endif
    return 0;
}

int orc_voxelize_slice(const float *vertices, uint32_t nv, const uint32_t *indices, uint32_t nt,
                       const uint8_t *materials, int mode, double param, int node_kind,
                       int direction, uint32_t slice, const uint32_t vox_split[2],
                       const uint32_t mat_split[2], void *nodes, uint32_t dims[3]) {
    if (node_kind != VOX_NODE_SHORT && node_kind != VOX_NODE_LONG && node_kind != VOX_NODE_PARTIAL)
        return 1;
    if (direction < 0 || direction > 2) return 2;
    // voxelizeSlice (src/host_side.cpp:1928-2120): voxSplitRes and
    // matSplitRes are (x, y) and y serves z too.
    const u32 vs[3] = {vox_split[0], vox_split[1], vox_split[1]};
    const u32 ms[3] = {mat_split[0], mat_split[1], mat_split[1]};
    std::vector<float> rot;
    const float *V = vertices;
    if (direction == 0) {  // voxelizeEntry rotates the model (703-716)
        rot = rotate_vertices(vertices, nv);
        V = rot.data();
    }
    vox_grid g;
    if (orc_plan_grid(V, nv, mode, param, vs[0], &g)) return 3;
    // The slice (voxelizeEntry, 743-780): y for directions 0 and 1, z for 2;
    // slice 0 and slices past the grid give zeros (sliceOOB).
    const int a = direction == 2 ? 1 : 0;
    bool oob = false;
    u32 s = slice;
    if (s == 0 || s > g.res[1 + a]) {
        s = 0;
        oob = true;
    } else {
        s -= 1;
    }
    u32 pmin[2] = {0, 0}, pmax[2] = {g.res[1], g.res[2]};
    pmin[a] = s;
    pmax[a] = s + 1;
    const PartLayout L = part_layout(g, pmin, pmax);
    const u32 X = g.alloc[0], Y = L.Yalloc, Z = L.Zalloc;
    std::vector<NodeRec> n =
        part_nodes(V, indices, nt, materials, g, L, vs, ms, node_kind == VOX_NODE_PARTIAL);
    if (direction == 0) {
        // restoreRotatedNodes, then erosion on the un-rotated array of
        // (Y, X, Z) nodes (1144-1220)
        n = unrotate_nodes(n, X, Y, Z);
        classify_part(n, Y, X, Z, vs, L.any_neighbour, false);
        dims[0] = Y;
        dims[1] = X;
    } else {
        classify_part(n, X, Y, Z, vs, L.any_neighbour, false);
        dims[0] = X;
        dims[1] = Y;
    }
    dims[2] = Z;
    if (oob)
        for (auto &r : n) r = {0, 0, 0.0f};
    if (nodes) pack_nodes(n, node_kind, nodes);
    return 0;
}

float orc_cut_volume(const float tri[9], const int32_t voxel[3], const float normal[3],
                     const float m[3], float d) {
    const F3 t[3] = {{tri[0], tri[1], tri[2]}, {tri[3], tri[4], tri[5]}, {tri[6], tri[7], tri[8]}};
    return cut_volume(t, voxel, F3{normal[0], normal[1], normal[2]}, F3{m[0], m[1], m[2]}, d);
}

uint32_t orc_tile_overlaps(const float tri[9], const double ext_min[3], double d, uint32_t ext_y,
                           uint32_t ext_z) {
    D3 v[3];
    for (int k = 0; k < 3; k++) v[k] = {tri[3 * k], tri[3 * k + 1], tri[3 * k + 2]};
    TileSetup ts;
    const u32 smin[2] = {0, 0}, smax[2] = {ext_y, ext_z};
    if (!tile_setup(v, ext_min, d, smin, smax, ts)) return 0;
    u32 n = 0;
    traverse_tiles(ts, ext_min, d, [&](u32, u32) { n++; });
    return n;
}

}  // extern "C"
