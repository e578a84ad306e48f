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
constexpr uint32_t kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

// Workgroups are dispatched round-robin over the XCDs.  Row kernels launch
// kXcds * ceil(n / kXcds) workgroups and map workgroup i to logical block
// (i % kXcds) * per + i / kXcds, so each XCD walks a contiguous run of rows
// and planes and its L2 serves the neighbouring-row reads.  Returns
// UINT32_MAX for the padding workgroups.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t n) {
    const uint32_t per = (n + kXcds - 1) / kXcds;
    const uint32_t L = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    return L < n ? L : UINT32_MAX;
}
inline uint32_t xcd_grid(uint32_t n) { return kXcds * ((n + kXcds - 1) / kXcds); }

// Logical row block -> (word chunk, 4-row group, plane).
struct RowBlock {
    uint32_t xc, y4, z;
};
__device__ __forceinline__ RowBlock row_block(uint32_t L, uint32_t nxc, uint32_t ny4) {
    RowBlock r;
    r.xc = L % nxc;
    const uint32_t q = L / nxc;
    r.y4 = q % ny4;
    r.z = q / ny4;
    return r;
}

// Surface bricks: default shapes of the Cartesian grid and of each FCC
// sub-grid; the shape is a run-time choice (the autotuner picks among
// power-of-two shapes with x a multiple of 64, at most 32 rows and at most
// kMaxBrickVox voxels).
constexpr int kCartBX = 128, kCartBY = 4, kCartBZ = 8;
constexpr int kFccBX = 64, kFccBY = 4, kFccBZ = 4;
constexpr int kMaxBrickVox = 4096;

// Brick geometry as the node passes need it: bricks per row/column and the
// log2 of the brick's words per row (x), rows (y) and planes (z).
struct BrickDims {
    uint32_t nbx, nby;
    uint32_t lwx, lby, lbz;
};
// z planes per thread in the erosion pass (a block covers 64 words x 4 rows
// x kErodeZ planes).
constexpr uint32_t kErodeZ = 8;
// z planes per thread in the Cartesian node pass; divides kCartBZ so a
// thread's planes share one surface brick.
constexpr uint32_t kNodeZ = 4;
static_assert(kCartBZ % kNodeZ == 0, "node planes must not straddle a brick");  // z >= 4
// Material-list arenas of the surface stage: kMatSlots counters, each on its
// own 128-byte line (kSlotStride words): {bytes used, surface voxels}.
constexpr uint32_t kMatSlots = 64, kSlotStride = 32;
// Cost model: features per brick, MLP parameters (16-64-64-1), sort buckets.
constexpr int kCostFeatures = 16;
constexpr int kCostParams = 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1;  // 5313
constexpr uint32_t kCostBuckets = 1024;
constexpr uint32_t kMlpBlockRows = 32;  // samples per MLP workgroup
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

// Block-level grouping of keys in an LDS hash table (256-thread blocks):
// candidates that land in the same bin share one global atomic per block.
constexpr uint32_t kAggSlots = 1024;
constexpr uint32_t kAggEmpty = 0xFFFFFFFFu;
struct AggTable {
    uint32_t key[kAggSlots];
    uint32_t cnt[kAggSlots];
    uint32_t base[kAggSlots];
};
__device__ __forceinline__ void agg_clear(AggTable &T) {
    for (uint32_t i = threadIdx.x; i < kAggSlots; i += blockDim.x) {
        T.key[i] = kAggEmpty;
        T.cnt[i] = 0u;
    }
}
// Inserts key; returns its slot and, in rank, its 0-based order among the
// block's items with the same key.
__device__ __forceinline__ uint32_t agg_insert(AggTable &T, uint32_t key, uint32_t &rank) {
    uint32_t h = (key * 2654435761u) >> 22;  // 10 bits
    while (true) {
        const uint32_t prev = atomicCAS(&T.key[h], kAggEmpty, key);
        if (prev == kAggEmpty || prev == key) break;
        h = (h + 1u) & (kAggSlots - 1u);
    }
    rank = atomicAdd(&T.cnt[h], 1u);
    return h;
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
    const float4 *rec;  // per-triangle vertex records (k_tri_records)
    uint32_t nt;
    double emx, emy, emz, d;
    uint32_t smin[2], smax[2];  // y/z subspace, ext coords
    uint32_t tilesY;            // (ext y + 3) / 4
    uint32_t X;                 // alloc x
    uint32_t rowsY;             // alloc y (= ext y + 2)
};

// Per-triangle records: vertices v0, v1, v2 in 48 contiguous bytes
// (rec[3t] = v0.xyz v1.x, rec[3t+1] = v1.yz v2.xy, rec[3t+2] = v2.z, class
// word (surface stage), 0, 0).
__global__ void k_tri_records(const float *__restrict__ V, const uint32_t *__restrict__ I,
                              uint32_t nt, float4 *__restrict__ rec);
__global__ void k_solid_setup(SolidParams p, uint32_t *__restrict__ cand_count,
                              uint4 *__restrict__ cbox);
__global__ void k_solid_candidates(SolidParams p, const uint32_t *__restrict__ cand_off,
                                   const uint32_t *__restrict__ cand_total,
                                   const uint4 *__restrict__ cbox, uint2 *__restrict__ pairs,
                                   uint32_t *__restrict__ tile_hist);
__global__ void k_solid_scatter(const uint2 *__restrict__ pairs,
                                const uint32_t *__restrict__ cand_total,
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
                             uint32_t *__restrict__ cand_count, uint4 *__restrict__ tri_ext,
                             float4 *__restrict__ tri_rec);
__global__ void k_surf_bin(SurfParams p, const uint32_t *__restrict__ cand_off,
                           const uint32_t *__restrict__ cand_total,
                           const uint4 *__restrict__ tri_info,
                           uint32_t *__restrict__ brick_hist, const uint32_t *__restrict__ brick_off,
                           uint32_t *__restrict__ brick_cursor, uint32_t *__restrict__ blist,
                           int scatter);
__global__ void k_surf_brick(SurfParams p, const SplitTable *__restrict__ splits,
                             const float4 *__restrict__ tri_rec,
                             const uint32_t *__restrict__ brick_off,
                             const uint32_t *__restrict__ brick_hist,
                             const uint32_t *__restrict__ blist,
                             const uint8_t *__restrict__ materials, uint32_t *__restrict__ bits,
                             uint32_t *__restrict__ surf, uint32_t *__restrict__ brick_mat_off,
                             uint16_t *__restrict__ wrank, uint8_t *__restrict__ mat_list,
                             uint32_t *__restrict__ slots, uint32_t slot_cap,
                             uint32_t *__restrict__ overflow, const uint32_t *__restrict__ order,
                             const uint32_t *__restrict__ n_order, uint32_t *__restrict__ next,
                             uint64_t *__restrict__ cycles);
__global__ void k_nonzero_flags(const uint32_t *__restrict__ hist, uint32_t *__restrict__ flags,
                                uint32_t n);
__global__ void k_nonzero_list(const uint32_t *__restrict__ hist, const uint32_t *__restrict__ off,
                               uint32_t *__restrict__ list, uint32_t n);

// ---------------------------------------------------------------------------
// Nodes
// ---------------------------------------------------------------------------
struct GridDims {
    uint32_t X, Y, Z;  // alloc dims of the (per-grid) bit array
    uint32_t W;        // X / 32
};

__global__ void k_erode(GridDims g, uint32_t *__restrict__ bits, uint32_t y0, uint32_t y1,
                        uint32_t z0, uint32_t z1, uint32_t *__restrict__ stamp,
                        uint32_t *__restrict__ changes, uint32_t pass, uint32_t gpass, int full);

// Boundary-id table (permutation -> id) in constant memory.
hipError_t set_orient_table(const uint8_t lut[64]);

struct MatSource {
    const uint32_t *surf;          // surface bit words (same layout as bits)
    const uint32_t *brick_mat_off; // per brick offset into mat_list or UINT32_MAX
    const uint16_t *wrank;         // per brick word: byte offset of its material segment in the brick
    const uint8_t *mat_list;
};

__global__ void k_nodes_cart(GridDims g, const uint32_t *__restrict__ solid,
                             const uint32_t *__restrict__ pre, const uint8_t *__restrict__ ystart,
                             const uint8_t *__restrict__ zstart, MatSource ms, int use_mats,
                             int kind, void *__restrict__ nodes, BrickDims bd);
__global__ void k_nodes_fcc(GridDims g, const uint32_t *s1, const uint32_t *s2,
                            const uint32_t *s3, const uint32_t *s4, MatSource m1, MatSource m2,
                            MatSource m3, MatSource m4, int use_mats, int kind,
                            void *__restrict__ nodes, BrickDims bd);

// Cost model kernels (vox_cost.hip).
__global__ void k_brick_features(const uint32_t *__restrict__ order,
                                 const uint32_t *__restrict__ sel,
                                 const uint32_t *__restrict__ n_rows,
                                 const uint32_t *__restrict__ brick_hist,
                                 const uint32_t *__restrict__ brick_off,
                                 const uint32_t *__restrict__ blist,
                                 const uint4 *__restrict__ tri_info,
                                 const uint4 *__restrict__ tri_ext, int32_t bd0, int32_t bd1,
                                 int32_t bd2, float *__restrict__ feats);
__global__ void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, uint32_t n,
                              float *__restrict__ Y);
__global__ void k_mlp_forward_n(const float *__restrict__ P, const float *__restrict__ X,
                                const uint32_t *__restrict__ n_dev, float *__restrict__ Y);
__global__ void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                           const float *__restrict__ T, uint32_t n, float inv_scale,
                           float *__restrict__ partial, float *__restrict__ ploss);
__global__ void k_mlp_reduce(const float *__restrict__ partial, const float *__restrict__ ploss,
                             uint32_t nblocks, float *__restrict__ grad, float *__restrict__ loss);
__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                       float *__restrict__ V, uint32_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2);
__global__ void k_cost_screen(const uint32_t *__restrict__ order,
                              const uint32_t *__restrict__ n_order,
                              const uint32_t *__restrict__ brick_hist,
                              const uint32_t *__restrict__ total_pairs, float ratio,
                              uint32_t *__restrict__ flags, uint32_t *__restrict__ light);
__global__ void k_cost_compact(const uint32_t *__restrict__ flags, const uint32_t *__restrict__ off,
                               const uint32_t *__restrict__ n_order, uint32_t *__restrict__ sel);
__global__ void k_cost_hist(const float *__restrict__ Y, const uint32_t *__restrict__ sel,
                            const uint32_t *__restrict__ n_sel, float heavy,
                            uint32_t *__restrict__ light, uint32_t *__restrict__ hist);
__global__ void k_cost_scatter(const float *__restrict__ Y, const uint32_t *__restrict__ order,
                               const uint32_t *__restrict__ n_order,
                               const uint32_t *__restrict__ sel, const uint32_t *__restrict__ n_sel,
                               float heavy, const uint32_t *__restrict__ off,
                               const uint32_t *__restrict__ light,
                               const uint32_t *__restrict__ loff, uint32_t *__restrict__ cursor,
                               uint32_t *__restrict__ out);

}  // namespace vk
