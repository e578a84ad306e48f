// vox_kernels.h -- HIP kernels of the voxelizer hot path (gfx950, wave64).
//
// Solid stage:   k_solid_setup -> scan -> k_solid_candidates -> scan ->
//                k_solid_scatter -> k_solid_fill (LDS-staged crossings +
//                per-row prefix XOR, one coalesced write per bit word).
// Surface stage: k_surf_setup -> scan -> k_surf_bin (count) -> scan ->
//                k_surf_bin (scatter) -> k_surf_brick (LDS atomicMax winners).
// Nodes:         k_erode (bitwise 6-neighbour erosion passes) ->
//                k_nodes_cart / k_nodes_fcc (bid/mask + material, node write).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vox_math.h"

namespace vk {

constexpr int kWave = 64;

// Cartesian bricks (node pass + surface) and FCC per-grid bricks.
constexpr int kCartBX = 128, kCartBY = 4, kCartBZ = 8;
constexpr int kFccBX = 64, kFccBY = 4, kFccBZ = 4;
constexpr int kMaxBrickVox = 4096;
constexpr int kMaxSplits = 64;

struct SplitTable {
    // matSplitRes subspace intervals per axis, ext coordinates, [min, max).
    uint32_t n[3];
    uint32_t lo[3][kMaxSplits];
    uint32_t hi[3][kMaxSplits];
};

struct SurfParams {
    const float *V;
    const uint32_t *I;
    uint32_t nt;
    F3 m;       // float(extMinVertex)
    float d;    // float(voxelLength)
    int32_t res[3];  // ext resolution (x, y, z)
    uint32_t alloc[3];
    int32_t bdim[3];  // brick dims in voxels
    uint32_t nb[3];   // bricks per axis
};

// ---------------------------------------------------------------------------
// Wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated atomic bump: one global atomic per wave.
__device__ __forceinline__ uint32_t wave_bump(uint32_t *counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    uint32_t lane = lane_id();
    uint32_t n = (uint32_t)__popcll(mask);
    uint32_t leader = n ? (uint32_t)(__ffsll((long long)mask) - 1) : 0u;
    uint32_t base = 0;
    if (n && lane == leader) base = atomicAdd(counter, n);
    base = __shfl(base, (int)leader);
    uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    return base + below;
}

__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *__restrict__ a, uint32_t n,
                                                    uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Scan (exclusive, uint32) -- three passes, 256 threads x 16 items per block.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_scan_blocks(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                              uint32_t n, uint32_t *__restrict__ partials);
__global__ void k_scan_add(uint32_t *__restrict__ out, uint32_t n,
                           const uint32_t *__restrict__ partials);

// ---------------------------------------------------------------------------
// Grid planning
// ---------------------------------------------------------------------------
__global__ void k_bbox(const float *__restrict__ V, uint32_t nv, uint32_t *__restrict__ out6);

// ---------------------------------------------------------------------------
// Solid
// ---------------------------------------------------------------------------
struct SolidParams {
    const float *V;
    const uint32_t *I;
    uint32_t nt;
    double emx, emy, emz, d;
    uint32_t smin[2], smax[2];  // y/z subspace, ext coords
    uint32_t tilesY;            // (ext y + 3) / 4
    uint32_t X;                 // alloc x
    uint32_t rowsY;             // alloc y (= ext y + 2)
};

__global__ void k_solid_setup(SolidParams p, uint32_t *__restrict__ cand_count);
__global__ void k_solid_candidates(SolidParams p, const uint32_t *__restrict__ cand_off,
                                   const uint32_t *__restrict__ cand_total,
                                   uint2 *__restrict__ pairs, uint32_t *__restrict__ pair_count,
                                   uint32_t pair_cap, uint32_t *__restrict__ tile_hist);
__global__ void k_solid_scatter(const uint2 *__restrict__ pairs,
                                const uint32_t *__restrict__ pair_count,
                                const uint32_t *__restrict__ tile_off,
                                uint32_t *__restrict__ tile_cursor, uint32_t *__restrict__ list);
__global__ void k_solid_fill(SolidParams p, const uint32_t *__restrict__ tile_off,
                             const uint32_t *__restrict__ tile_hist,
                             const uint32_t *__restrict__ list, uint32_t ty0, uint32_t nty,
                             uint32_t tz0, uint32_t *__restrict__ bits);
__global__ void k_zero_pad_rows(uint32_t *__restrict__ bits, uint32_t words_per_row,
                                uint32_t rowsY, uint32_t rowsZ);

// ---------------------------------------------------------------------------
// Surface
// ---------------------------------------------------------------------------
__global__ void k_surf_setup(SurfParams p, uint4 *__restrict__ tri_info,
                             uint32_t *__restrict__ cand_count);
__global__ void k_surf_bin(SurfParams p, const uint32_t *__restrict__ cand_off,
                           const uint32_t *__restrict__ cand_total,
                           const uint4 *__restrict__ tri_info,
                           uint32_t *__restrict__ brick_hist, const uint32_t *__restrict__ brick_off,
                           uint32_t *__restrict__ brick_cursor, uint32_t *__restrict__ blist,
                           int scatter);
__global__ void k_surf_brick(SurfParams p, const SplitTable *__restrict__ splits,
                             const uint4 *__restrict__ tri_info,
                             const uint32_t *__restrict__ brick_off,
                             const uint32_t *__restrict__ brick_hist,
                             const uint32_t *__restrict__ blist,
                             const uint8_t *__restrict__ materials, uint32_t *__restrict__ bits,
                             uint32_t *__restrict__ surf, uint32_t *__restrict__ brick_mat_off,
                             uint8_t *__restrict__ mat_list, uint32_t *__restrict__ mat_cursor,
                             uint32_t mat_cap, uint32_t *__restrict__ stats);

// ---------------------------------------------------------------------------
// Nodes
// ---------------------------------------------------------------------------
struct GridDims {
    uint32_t X, Y, Z;  // alloc dims of the (per-grid) bit array
    uint32_t W;        // X / 32
};

__global__ void k_erode(GridDims g, const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                        uint32_t y0, uint32_t y1, uint32_t z0, uint32_t z1,
                        uint32_t *__restrict__ changes, uint32_t pass);

// Boundary-id table (permutation -> id) in constant memory.
hipError_t set_orient_table(const uint8_t lut[64]);

struct MatSource {
    const uint32_t *surf;          // surface bit words (same layout as bits)
    const uint32_t *brick_mat_off; // per brick offset into mat_list or UINT32_MAX
    const uint8_t *mat_list;
};

__global__ void k_nodes_cart(GridDims g, const uint32_t *__restrict__ solid,
                             const uint32_t *__restrict__ pre, const uint8_t *__restrict__ ystart,
                             const uint8_t *__restrict__ zstart, MatSource ms, int use_mats,
                             int kind, void *__restrict__ nodes, uint32_t nbx, uint32_t nby);
__global__ void k_nodes_fcc(GridDims g, const uint32_t *s1, const uint32_t *s2,
                            const uint32_t *s3, const uint32_t *s4, MatSource m1, MatSource m2,
                            MatSource m3, MatSource m4, int use_mats, int kind,
                            void *__restrict__ nodes, uint32_t nbx, uint32_t nby);

}  // namespace vk
