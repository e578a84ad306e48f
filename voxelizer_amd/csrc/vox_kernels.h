// vox_kernels.h -- HIP kernels of the voxelizer hot path (gfx950, wave64).
//
// Solid stage:   k_solid_setup -> scan -> k_block_starts -> k_solid_candidates -> scan ->
//                k_solid_scatter -> k_solid_fill (LDS-staged crossings +
//                per-row prefix XOR, one coalesced write per bit word).
// Surface stage: k_surf_setup -> scan -> k_block_starts -> k_surf_bin (count) -> scan ->
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

// Global-memory view of a pointer read from a job table.  Pointers the
// compiler cannot place in an address space are accessed with flat_*
// instructions, which also count against the LDS wait counter, so every
// later LDS read waits for them; global_* accesses do not.
#define VK_G __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ VK_G T *gptr(T *p) {
    return (VK_G T *)p;
}
// Native 4 x u32 vector for 16-byte global accesses through VK_G pointers.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

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

// Plane-linear blocks of the erosion and Cartesian node passes: the words of
// a plane are numbered i = y * W + w and cut into chunks of kChunk
// consecutive words; a 256-thread block takes one chunk over a group of
// planes, so every lane has a word whatever the row width W.
constexpr uint32_t kChunk = 256;
// x words of a solid-fill pass: a tile's 16 rows of them fill 64 KB of LDS;
// wider rows take several passes.
constexpr uint32_t kFillSegWords = 1024;
__host__ __device__ __forceinline__ uint32_t plane_chunks(uint32_t W, uint32_t Y) {
    return (W * Y + kChunk - 1u) / kChunk;
}

// Surface bricks: default shapes of the Cartesian grid and of each FCC
// sub-grid; the shape is a run-time choice (the autotuner picks among
// power-of-two shapes with x a multiple of 64, at most 32 rows and 256 to
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
struct GridDims {
    uint32_t X, Y, Z;  // alloc dims of the (per-grid) bit array
    uint32_t W;        // X / 32
};
// z planes per thread in the erosion pass (a block covers kChunk words of a
// plane x kErodeZ planes).
constexpr uint32_t kErodeZ = 8;
// z planes per thread in the Cartesian node pass; divides kCartBZ so a
// thread's planes share one surface brick.
constexpr uint32_t kNodeZ = 4;
static_assert(kCartBZ % kNodeZ == 0, "node planes must not straddle a brick");  // z >= 4
// Cost model: features per brick, MLP parameters (16-64-64-1).
constexpr int kCostFeatures = 16;
constexpr int kCostParams = 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1;  // 5313
constexpr uint32_t kMlpBlockRows = 32;  // samples per MLP workgroup
constexpr int kMaxSplits = 64;

struct SplitTable {
    // matSplitRes subspace intervals per axis, ext coordinates, [min, max).
    uint32_t n[3];
    uint32_t lo[3][kMaxSplits];
    uint32_t hi[3][kMaxSplits];
};

struct SurfParams {
    F3 m;       // float(extMinVertex)
    float d;    // float(voxelLength)
    int32_t res[3];  // ext resolution (x, y, z)
    uint32_t alloc[3];
    int32_t bdim[3];  // brick dims in voxels
    uint32_t nb[3];   // bricks per axis
    int32_t off[3];   // alloc coordinate of ext coordinate 0 (1, Job::oy, Job::oz)
};

// One grid to voxelize: a mesh on a Cartesian grid, or one of the four
// shifted sub-grids of an FCC lattice.  Any number of jobs share each stage's
// launches: the stage's work items (triangle slots, tiles, bricks, row
// blocks) are numbered over all jobs, and a job owns the contiguous run that
// starts at its *_base.  A job's triangle slots start at a multiple of 256,
// so every 256-thread block of a per-triangle pass belongs to one job.
struct Job {
    const float *V;        // mesh vertices (x, y, z)
    const uint32_t *I;     // mesh triangles (3 vertex indices)
    const uint8_t *mats;   // material per mesh triangle, or null
    uint32_t nt;           // mesh triangles
    uint32_t tri_base;     // first triangle slot
    GridDims gd;           // alloc dims of the bit grids
    int32_t res[3];        // ext resolution
    uint32_t *bits;        // solid bit words
    // solid stage
    double em[3];          // grid origin in FP64 (FCC sub-grids: shifted)
    double d;              // voxel length
    uint32_t tilesY, tilesZ;  // 4x4 tiles of voxel-centre rows
    uint32_t tile_base;    // first tile (one fill workgroup per tile)
    uint32_t pad_base;     // first block of the pad-row clear
    // Alloc layout of the ext space (a partition of a grid, as the
    // reference's devConfig / slice subspaces, src/host_side.cpp:901-1011):
    // alloc row / plane of ext row / plane 0 (1 behind a padding row, 0 when
    // the first ext row is the overlap row of a neighbouring partition).
    uint32_t oy, oz;
    uint32_t padmask;      // alloc rows / planes that are padding: 1 y = 0, 2 y = Y-1, 4 z = 0, 8 z = Z-1
    int32_t zero_border;   // clear the node array's border after classification
    int32_t rot_ids;       // ids of a model rotated about z (x slices): -x/+x <- -y/+y, -y/+y <- +x/-x
    // voxSplitRes y/z subspaces of the solid stage when there are several:
    // sstart[y] (y < res[1]) and sstart[res[1] + z] (z < res[2]) are 1 where
    // a subspace other than the first starts; null for one subspace.
    const uint8_t *sstart;
    // surface stage
    F3 m;                  // float(origin)
    float fd;              // float(voxel length)
    int32_t bdim[3];       // brick dims (voxels)
    uint32_t nb[3];        // bricks per axis
    uint32_t brick_base;   // first brick
    BrickDims bdn;         // brick geometry for the node passes
    const SplitTable *splits;  // matSplitRes intervals
    uint32_t *surf;        // surface bit words (the layout of bits); FCC and PartialNode only
    const uint32_t *bhist; // per brick: (triangle, brick) pairs; 0 = not visited
    uint32_t *wmask;       // per brick, 4 words: bit i set where brick word i holds a surface voxel
    uint8_t *matgrid;      // per voxel, brick-major: winner material (surface words)
    int32_t use_mats;      // surface stage and materials wanted
    int32_t embed;         // winner keys carry the material (mesh-local index < 2^22)
    // erosion and nodes (Cartesian jobs; FCC sub-grids have no such blocks)
    uint32_t erode_base, node_base;
    void *nodes;
    int32_t kind;          // 0 ShortNode, 1 LongNode, 3 ShortFCCNode, 4 LongFCCNode
    // PartialNode grids run as LongNode (nodes: a LongNode array) and are
    // expanded into pnodes afterwards; keygrid holds, laid out as matgrid
    // with 4 bytes per voxel, the winner key of every voxel of the surface
    // words.
    uint32_t *keygrid;
    void *pnodes;
    uint32_t *pbits;       // PartialNode: the solid|surface bits before erosion
};

// Job owning item x of a pass whose jobs start at bases[0..n) (ascending;
// jobs with no items share the next job's base): the last job whose base is
// <= x.  Uniform when x is.
__device__ __forceinline__ uint32_t job_of(const uint32_t *__restrict__ bases, uint32_t n,
                                           uint32_t x) {
    if (n == 1) return 0u;
    uint32_t lo = 0, hi = n;
    while (lo + 1u < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bases[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ SurfParams surf_params(const Job &J) {
    SurfParams p;
    p.m = J.m;
    p.d = J.fd;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        p.res[a] = J.res[a];
        p.bdim[a] = J.bdim[a];
        p.nb[a] = J.nb[a];
    }
    p.alloc[0] = J.gd.X;
    p.alloc[1] = J.gd.Y;
    p.alloc[2] = J.gd.Z;
    p.off[0] = 1;
    p.off[1] = (int32_t)J.oy;
    p.off[2] = (int32_t)J.oz;
    return p;
}

// ---------------------------------------------------------------------------
// Wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Where lanes read LDS that other lanes of their wave wrote: the hardware
// keeps a wave's LDS operations in order, but the compiler may move a lane's
// access past another to an address that lane itself never shares.  The
// wavefront-scope fences around the barrier pin the order.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive sum over the wave's lanes in registers (DPP; every lane active):
// shifts by 1, 2, 4 and 8 within each 16-lane row, then each row's total
// broadcast into the rows above it.  A lane whose source lies outside its
// row, or whose row is masked off, adds 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118,
                  kRowBcast15 = 0x142, kRowBcast31 = 0x143;
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr1, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr2, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr4, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowShr8, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast15, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kRowBcast31, 0xC, 0xF, false);
    return v;
}

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
// Scan (exclusive, uint32), 256 threads x 16 items per tile.  Up to
// kScanDirectTiles tiles: k_scan_reduce (tile sums) then k_scan_down (each
// tile adds the sums before it and scans itself); more: tiles scanned with
// k_scan_blocks, their sums scanned recursively, then k_scan_add.  The
// output has n + 2 entries: out[n] is the total mod 2^32 and out[n + 1] is 1
// when the true total reached 2^32 (the offsets wrapped), else 0.  partials
// holds 2 * tiles + 1 words (tile sums' low words, then their high words).
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kScanDirectTiles = 4096;

__global__ void k_scan_reduce(const uint32_t *__restrict__ in, uint32_t n,
                              uint32_t *__restrict__ partials);
__global__ void k_scan_down(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                            uint32_t n, const uint32_t *__restrict__ partials);

__global__ void k_scan_blocks(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                              uint32_t n, uint32_t *__restrict__ partials);
__global__ void k_scan_add(uint32_t *__restrict__ out, uint32_t n,
                           const uint32_t *__restrict__ partials,
                           const uint32_t *__restrict__ psum);

// ---------------------------------------------------------------------------
// Grid planning
// ---------------------------------------------------------------------------
__global__ void k_bbox(const float *__restrict__ V, uint32_t nv, uint32_t *__restrict__ out6);

// ---------------------------------------------------------------------------
// Solid
// ---------------------------------------------------------------------------
struct SolidParams {
    const float4 *rec;  // per-triangle vertex records (k_tri_records)
    double emx, emy, emz, d;
    uint32_t smin[2], smax[2];  // y/z extent of the centres, ext coords
    uint32_t tilesY;            // (ext y + 3) / 4
    uint32_t X;                 // alloc x
    uint32_t rowsY;             // alloc y
    uint32_t oy, oz;            // alloc row / plane of ext row / plane 0
};
__device__ __forceinline__ SolidParams solid_params(const Job &J, const float4 *rec) {
    SolidParams p;
    p.rec = rec;
    p.emx = J.em[0];
    p.emy = J.em[1];
    p.emz = J.em[2];
    p.d = J.d;
    p.smin[0] = 0u;
    p.smin[1] = 0u;
    p.smax[0] = (uint32_t)J.res[1];
    p.smax[1] = (uint32_t)J.res[2];
    p.tilesY = J.tilesY;
    p.X = J.gd.X;
    p.rowsY = J.gd.Y;
    p.oy = J.oy;
    p.oz = J.oz;
    return p;
}

// The (triangle, tile) and (triangle, brick) pair passes take the first
// min(total, cap) candidates; a total above cap sets *ovf (a replayed graph
// sized by an earlier run reports that its lists were too short).
// Per-triangle records: vertices v0, v1, v2 in 48 contiguous bytes
// (rec[3t] = v0.xyz v1.x, rec[3t+1] = v1.yz v2.xy, rec[3t+2] = v2.z, class
// word (surface stage), 0, 0), one per triangle slot of every job; tri_job
// holds each slot's job.  nslots is a multiple of 256 and the per-triangle
// kernels run 256-thread blocks.
__global__ void k_tri_records(const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_bases,
                              uint32_t njobs, uint32_t nslots, float4 *__restrict__ rec,
                              uint32_t *__restrict__ tri_job);
__global__ void k_solid_setup(const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_bases,
                              uint32_t njobs, uint32_t nslots, const float4 *__restrict__ rec,
                              uint32_t *__restrict__ cand_count, uint4 *__restrict__ cbox);
// first[b] = the triangle holding candidate 256*b of a candidate pass (cand_off:
// exclusive offsets, nslots + 1 entries); candidates at or past cap are skipped.
__global__ void k_block_starts(const uint32_t *__restrict__ cand_off, uint32_t nslots,
                               uint32_t cap, uint32_t *__restrict__ first);
inline size_t block_starts_size(uint32_t total) { return (size_t)(total + 255u) / 256u + 1u; }
// With pmask (jobs with several y/z subspaces), pmask[c] receives the rows of
// the candidate's tile whose subspace queues the triangle (bit r = centre row
// r of the tile, y fastest); the scatter pass carries it into list_mask and
// the fill applies a triangle's crossings only on those rows.
template <bool kMulti>
__global__ void k_solid_candidates(const Job *__restrict__ jobs,
                                   const uint32_t *__restrict__ tri_job, uint32_t nslots,
                                   const float4 *__restrict__ rec,
                                   const uint32_t *__restrict__ cand_off,
                                   const uint32_t *__restrict__ cand_total,
                                   const uint32_t *__restrict__ first,
                                   const uint4 *__restrict__ cbox, uint2 *__restrict__ pairs,
                                   uint16_t *__restrict__ pmask,
                                   uint32_t *__restrict__ tile_hist, uint32_t cap,
                                   uint32_t *__restrict__ ovf);
__global__ void k_solid_scatter(const uint2 *__restrict__ pairs,
                                const uint16_t *__restrict__ pmask,
                                const uint32_t *__restrict__ cand_total,
                                const uint32_t *__restrict__ tile_off,
                                uint32_t *__restrict__ tile_cursor, uint32_t *__restrict__ list,
                                uint16_t *__restrict__ list_mask, uint32_t cap);
// One workgroup per tile of every job (grid = all tiles).
__global__ void k_solid_fill(const Job *__restrict__ jobs, const uint32_t *__restrict__ tile_bases,
                             uint32_t njobs, const float4 *__restrict__ rec,
                             const uint32_t *__restrict__ tile_off,
                             const uint32_t *__restrict__ tile_hist,
                             const uint32_t *__restrict__ list,
                             const uint16_t *__restrict__ list_mask);
// Zeroes the pad rows (y or z at the alloc border); grid = all jobs' blocks.
__global__ void k_zero_pad_rows(const Job *__restrict__ jobs,
                                const uint32_t *__restrict__ pad_bases, uint32_t njobs);
__host__ __device__ __forceinline__ uint32_t pad_words(uint32_t W, uint32_t rowsY,
                                                    uint32_t rowsZ) {
    return (2u * rowsZ + 2u * rowsY) * W;
}

// ---------------------------------------------------------------------------
// Surface
// ---------------------------------------------------------------------------
__global__ void k_surf_setup(const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_bases,
                             uint32_t njobs, uint32_t nslots, uint4 *__restrict__ tri_info,
                             uint32_t *__restrict__ cand_count, uint4 *__restrict__ tri_ext,
                             float4 *__restrict__ tri_rec);
template <bool kMulti>
__global__ void k_surf_bin(const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_job,
                           uint32_t nslots, const uint32_t *__restrict__ cand_off,
                           const uint32_t *__restrict__ cand_total,
                           const uint32_t *__restrict__ first,
                           const uint4 *__restrict__ tri_info,
                           uint32_t *__restrict__ brick_hist, const uint32_t *__restrict__ brick_off,
                           uint32_t *__restrict__ brick_cursor, uint32_t *__restrict__ blist,
                           int scatter, uint32_t cap, uint32_t *__restrict__ ovf);
// kKeys: jobs with a keygrid also get their winner keys there (PartialNode).
template <bool kKeys>
__global__ void k_surf_brick(const Job *__restrict__ jobs, const uint32_t *__restrict__ brick_bases,
                             uint32_t njobs, const float4 *__restrict__ tri_rec,
                             const uint32_t *__restrict__ brick_off,
                             const uint32_t *__restrict__ brick_hist,
                             const uint32_t *__restrict__ blist, uint32_t *__restrict__ stats,
                             const uint32_t *__restrict__ order,
                             const uint32_t *__restrict__ n_order,
                             const uint32_t *__restrict__ hlist,
                             const uint32_t *__restrict__ hsub,
                             const uint32_t *__restrict__ n_heavy,
                             const uint8_t *__restrict__ skip, uint32_t *__restrict__ next,
                             uint64_t *__restrict__ cycles);
__global__ void k_nonzero_flags(const uint32_t *__restrict__ hist, uint32_t *__restrict__ flags,
                                uint32_t n);
__global__ void k_nonzero_list(const uint32_t *__restrict__ hist, const uint32_t *__restrict__ off,
                               uint32_t *__restrict__ list, uint32_t n);

// ---------------------------------------------------------------------------
// Nodes
// ---------------------------------------------------------------------------
// Erosion pass over every Cartesian job's row blocks (grid = all blocks);
// voxels outside [y0,y1) x [z0,z1) are left alone.  changes[pass * njobs + j]
// is set to 1 if pass `pass` changed job j, totals[pass] if it changed any,
// and *notconv too when given (the last pass of a fixed pass count).
__global__ void k_erode(const Job *__restrict__ jobs, const uint32_t *__restrict__ erode_bases,
                        uint32_t njobs, uint32_t y0, uint32_t y1, uint32_t z0, uint32_t z1,
                        uint32_t *__restrict__ stamp, uint32_t *__restrict__ changes,
                        uint32_t *__restrict__ totals, uint32_t *__restrict__ notconv,
                        uint32_t pass, uint32_t gpass, int full);
inline uint32_t erode_blocks(const GridDims &g) {
    return plane_chunks(g.W, g.Y) * ((g.Z + kErodeZ - 1u) / kErodeZ);
}

// Boundary ids, bit-sliced: bit b of the ids of 32 voxels from the words of
// their neighbour bits (mx/px: the -x/+x neighbour is solid; my/py, mz/pz
// likewise).  This is the permutation -> id table of
// src/voxelizer.cu:2142-2236 (ids documented at include/voxelizer.h:47-83;
// permutation bits -x 1, +x 2, -y 4, +y 8, -z 16, +z 32) as two-level logic
// over its 64 entries, minimised with the 37 permutations that leave an axis
// without neighbours as don't-cares and then masked to 0 there, as the
// table has them.  vox_create checks it against the table entry by entry.
struct IdPlanes {
    uint32_t b[5];
};
__host__ __device__ inline IdPlanes orient_planes(uint32_t mx, uint32_t px, uint32_t my,
                                                  uint32_t py, uint32_t mz, uint32_t pz) {
    const uint32_t valid = (mx | px) & (my | py) & (mz | pz);
    IdPlanes o;
    o.b[0] = valid & (~px | (mx & ~py) | (mx & my & mz));
    o.b[1] = valid & ((mx & ~my) | (~mx & ~py) | (mx & py & ~mz) | (mx & mz & px & pz) |
                      (py & ~px & ~pz));
    o.b[2] = valid & ((~mz & ~px) | (~mz & ~py) | (mz & ~mx & ~my) | (mx & px & pz & ~my) |
                      (mx & px & pz & ~py) | (my & px & py & ~pz));
    o.b[3] = valid & ((mx & px & ~mz) | (my & py & ~px) | (my & mz & py & pz) |
                      (my & mz & py & ~mx) | (~mx & ~my & ~mz) | (mx & px & ~my & ~pz) |
                      (mx & px & ~py & ~pz));
    o.b[4] = valid & ((mz & pz) | (mx & my & px & py) | (my & py & ~mx & ~mz));
    return o;
}

// Surface data of one grid as the node passes read it.  Brick b's voxels
// occupy bytes [b * nvox, (b + 1) * nvox) of matgrid in the brick's word
// order (z, y, x-word), 32 bytes per word.  A word that holds a surface
// voxel has all 32 bytes written in the call (0 for its other voxels); the
// rest may be stale.  Bricks the surface pass did not visit (bhist == 0)
// hold stale surface words, so the readers check bhist first.
struct MatSource {
    const VK_G uint32_t *surf;   // surface bit words (same layout as bits)
    const VK_G uint32_t *bhist;  // per brick: nonzero if visited
    const VK_G uint8_t *matgrid;
};
__device__ __forceinline__ MatSource mat_source(const Job &J) {
    return MatSource{gptr(J.surf), gptr(J.bhist), gptr(J.matgrid)};
}
// Index of word (w, y, z) of a grid within its brick's block of words, and
// the brick: brick-major, words in (z, y, x-word) order inside a brick.
__device__ __forceinline__ uint32_t brick_of_word(const BrickDims &bd, uint32_t w, uint32_t y,
                                                  uint32_t z) {
    return ((z >> bd.lbz) * bd.nby + (y >> bd.lby)) * bd.nbx + (w >> bd.lwx);
}
__device__ __forceinline__ uint32_t word_in_brick(const BrickDims &bd, uint32_t w, uint32_t y,
                                                  uint32_t z) {
    return (((z & ((1u << bd.lbz) - 1u)) << (bd.lby + bd.lwx)) +
            ((y & ((1u << bd.lby) - 1u)) << bd.lwx)) + (w & ((1u << bd.lwx) - 1u));
}

// Node passes.  Cartesian: every Cartesian job's row blocks (grid = all);
// pre/ystart/zstart only with a single job.  FCC: per FCC lattice f, its
// four sub-grids are jobs fcc_first[f] .. +3 (the first holds the node array
// and kind, 3 ShortFCCNode or 4 LongFCCNode) and its blocks start at
// fcc_bases[f].
__global__ void k_nodes_cart(const Job *__restrict__ jobs, const uint32_t *__restrict__ node_bases,
                             uint32_t njobs);
// The same with neighbours across subspace seams read from pre (ystart/zstart).
__global__ void k_nodes_cart_pre(const Job *__restrict__ jobs,
                                 const uint32_t *__restrict__ node_bases, uint32_t njobs,
                                 const uint32_t *__restrict__ pre,
                                 const uint8_t *__restrict__ ystart,
                                 const uint8_t *__restrict__ zstart);
// PartialNode expansion of job j (after the node pass): each LongNode of
// its nodes array becomes {bid, mat, 0, 0, ratio} in pnodes, the ratio 1 or 0
// for a solid or empty node and the winner triangle's cut ratio for a
// surface node (0 on a cleared partition border).
__global__ void k_nodes_partial(const Job *__restrict__ jobs, uint32_t j);
inline uint32_t node_blocks(const GridDims &g) {
    return plane_chunks(g.W, g.Y) * ((g.Z + kNodeZ - 1u) / kNodeZ);
}
__global__ void k_nodes_fcc(const Job *__restrict__ jobs, const uint32_t *__restrict__ fcc_first,
                            const uint32_t *__restrict__ fcc_bases, uint32_t nfcc);
inline uint32_t fcc_node_blocks(const GridDims &g) {
    return ((2u * g.W + kWave - 1u) / kWave) * ((g.Y + 3u) / 4u) * 2u * g.Z;
}

// Cost model kernels (vox_cost.hip).
__global__ void k_brick_features(const uint32_t *__restrict__ order,
                                 const uint32_t *__restrict__ sel,
                                 const uint32_t *__restrict__ n_rows,
                                 const uint32_t *__restrict__ brick_hist,
                                 const uint32_t *__restrict__ brick_off,
                                 const uint32_t *__restrict__ blist,
                                 const uint4 *__restrict__ tri_info,
                                 const uint4 *__restrict__ tri_ext, const Job *__restrict__ jobs,
                                 const uint32_t *__restrict__ brick_bases, uint32_t njobs,
                                 float *__restrict__ feats);
__global__ void k_mlp_forward(const float *__restrict__ P, const float *__restrict__ X, uint32_t n,
                              float *__restrict__ Y);
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
                              uint32_t *__restrict__ cand, uint32_t *__restrict__ cnt,
                              uint8_t *__restrict__ skip);
// Features, MLP forward and heavy test of the screened candidates in one
// launch (one wave per brick); the learned schedule's path.
__global__ void k_cost_predict(const float *__restrict__ P, const uint32_t *__restrict__ order,
                               const uint32_t *__restrict__ cand, uint32_t *__restrict__ cnt,
                               float heavy, uint32_t max_parts,
                               const uint32_t *__restrict__ brick_hist,
                               const uint32_t *__restrict__ brick_off,
                               const uint32_t *__restrict__ blist,
                               const uint4 *__restrict__ tri_info, const uint4 *__restrict__ tri_ext,
                               const Job *__restrict__ jobs,
                               const uint32_t *__restrict__ brick_bases, uint32_t njobs,
                               uint32_t *__restrict__ hlist, uint32_t *__restrict__ hsub,
                               uint8_t *__restrict__ skip);
// Standardized log cost, above the heavy threshold, per doubling of the
// number of sub-bricks a heavy brick is cut into (k_cost_predict).
constexpr float kSplitStep = 1.0f;
// Heavy-list entries per listed brick at most (sub-bricks).
constexpr uint32_t kMaxSplitParts = 4;

}  // namespace vk
