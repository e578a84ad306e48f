// vox_api.hip -- C ABI of libvoxelizer_amd.so: context, device arena,
// planning and the stage pipelines (host side, HIP runtime).
//
// Replaces the reference's host orchestration for one device
// (src/host_side.cpp:275-397 performVoxelization, 901-1244 voxelizeWorker,
// 1262-1698 fccWorker) and its launcher layer (src/voxelizer.cu:39-806).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/voxelizer_amd.h"
#include "vox_kernels.h"

using namespace vk;

namespace {

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
};

struct Span {
    uint32_t min, max;
};

// splitResolutionByNrOfParts for one axis (src/host_side.cpp:3003-3064).
std::vector<Span> split_axis(uint32_t len, uint32_t parts) {
    std::vector<Span> out(parts);
    uint32_t prev = len;
    for (uint32_t i = parts; i > 0; --i) {
        Span s;
        s.max = prev;
        s.min = s.max - s.max / i;
        prev = s.min;
        out[i - 1] = s;
    }
    return out;
}
std::vector<Span> split_axis_max(uint32_t len, uint32_t maxdim) {
    return split_axis(len, (len + maxdim - 1) / maxdim);
}

const uint8_t kOrientPairs[27][2] = {
    {21, 1},  {22, 2},  {25, 3},  {26, 4},  {37, 5},  {38, 6},  {41, 7},
    {42, 8},  {23, 9},  {27, 10}, {29, 11}, {30, 12}, {39, 13}, {43, 14},
    {45, 15}, {46, 16}, {53, 17}, {54, 18}, {57, 19}, {58, 20}, {31, 21},
    {59, 22}, {55, 23}, {62, 24}, {61, 25}, {47, 26}, {63, 27}};

// Scratch of vox_voxelize_nodes_noalloc: a file-scope device array, one per
// GPU the library runs on, held until the process ends.  Host code never
// uses the array's name as an address: each device's copy is looked up once
// with hipGetSymbolAddress (scratch_on) and that pointer is used instead.
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices];
bool g_scratch_done[kMaxScratchDevices];

// Device address of `scratch` on `device` (current), looked up once per
// device; null (and the HIP error in *err) if the lookup fails.
void *scratch_on(int device, hipError_t *err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            *err = e;
            return nullptr;
        }
        g_scratch_ptr[device] = ptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_ptr[device];
}

// Named regions of the scratch array during one no-allocation call (a fixed
// table: that call allocates nothing, not even host memory for names).
constexpr int kArenaSlots = 96;
struct ArenaSlot {
    char name[32];
    size_t off, cap;
};

}  // namespace

struct vox_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, Buf> bufs;
    // mesh
    const float *V = nullptr;
    const uint32_t *I = nullptr;
    uint32_t nv = 0, nt = 0;
    bool own_mesh = false;
    const uint8_t *mats = nullptr;
    bool own_mats = false;
    uint32_t n_unique = 0;
    // host staging for small read-backs
    uint32_t *pinned = nullptr;
    // timing
    bool timing = false;
    hipEvent_t ev[8] = {};
    vox_stage_times times = {};
    // persistent surface launch: compute units and resident bricks per CU
    int n_cu = 0;
    int surf_blocks_per_cu = 0;
    int erode_blocks_per_cu = 0;  // resident erosion workgroups per CU
    // brick schedule of the surface pass and its cost model
    int schedule = VOX_SCHEDULE_PLAIN;
    float *cost_params = nullptr;  // device, kCostParams floats
    bool cost_ready = false;
    float heavy_cost = 1.5f;    // standardized log cost above which a brick goes first
    // Pair count / mean from which a brick's cost is predicted.  Measured on
    // the large hall and the 256-mesh batch (Gvoxels/s): 6 -> 571 / 365,
    // 12 -> 576 / 366, 24 -> 579 / 368 in one run, 24 -> 577 / 362,
    // 48 -> 567 / 359, 96 -> 560 / 342 in another.  Below 24 the features of
    // light bricks cost more than their order gains; above it heavy bricks
    // are missed.
    float screen_ratio = 24.0f;
    // Sub-bricks per heavy brick at most (1, 2 or 4); 0: from the number of
    // bricks per resident brick-pass workgroup.
    uint32_t split_parts = 0;
    bool record = false;
    std::vector<float> rec_feats, rec_cycles;
    // worker contexts of batched runs (same device, own streams and buffers)
    std::vector<vox_ctx *> workers;
    // per-triangle-slot vertex records and job ids of the current call
    float4 *rec = nullptr;
    uint32_t *tri_job = nullptr;
    // pinned staging area of job-table uploads, and the event of the last one
    void *up = nullptr;
    size_t up_cap = 0;
    hipEvent_t up_ev = nullptr;
    // The surface stage's preparation (setup, binning, brick order) runs on
    // `side` beside the solid stage; fork and join through these events.
    hipStream_t side = nullptr, main_stream = nullptr;  // main_stream: `stream` as created
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // surface brick shapes (voxels along x, y, z)
    int cart_bdim[3] = {kCartBX, kCartBY, kCartBZ};
    // meshes with fewer triangles carry materials in the surface pass's
    // winner keys (at most 2^22; VOX_EMBED_TRIS lowers it, for tests of the
    // other path)
    uint32_t embed_limit = 1u << 22;
    // Fixed sizes (graph capture): the pair-list capacities and the number
    // of erosion passes come from an earlier run instead of read-backs; the
    // kernels raise ovf[0] when a list was too short, and last_changed points
    // at the change flag of the last erosion pass.  The last_* fields hold
    // the sizes the latest run needed.
    bool fixed = false;
    uint32_t cap_solid = 0, cap_surf = 0, n_passes = 0;
    uint32_t *ovf = nullptr;
    const uint32_t *last_changed = nullptr;
    uint32_t last_solid = 0, last_surf = 0, last_passes = 0;
    int fcc_bdim[3] = {kFccBX, kFccBY, kFccBZ};
    // No-allocation mode (vox_voxelize_nodes_noalloc): buffers are carved
    // from this device's scratch array (scratch_dev, set up by vox_create)
    // instead of hipMalloc.
    uint8_t *scratch_dev = nullptr;
    bool arena = false;
    size_t arena_used = 0;
    int arena_n = 0;
    ArenaSlot arena_slot[kArenaSlots];
};

#define VOX_HIP(expr)                                                               \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(hipGetErrorString(e_)) + " @ " + #expr;          \
            return e_ == hipErrorOutOfMemory ? VOX_ERR_OOM : VOX_ERR_HIP;           \
        }                                                                           \
    } while (0)

// For entry points whose context is g->w.
#define VOX_HIP_G(expr)                                        \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) {                                \
            g->w->err = std::string(hipGetErrorString(e_));    \
            return VOX_ERR_HIP;                                \
        }                                                      \
    } while (0)

#define VOX_TRY(expr)                 \
    do {                              \
        vox_status s_ = (expr);       \
        if (s_ != VOX_OK) return s_;  \
    } while (0)

namespace {

// Pinned staging buffer for small device-to-host reads (counters, totals).
constexpr size_t kPinnedBytes = 64 * 1024;

vox_status fail(vox_ctx *ctx, vox_status s, const std::string &msg) {
    ctx->err = msg;
    return s;
}

// A named region of the scratch array (no-allocation mode): reused when the
// name was given out before in this call and is large enough, else carved
// after the regions given out so far.
vox_status arena_buf(vox_ctx *ctx, const char *name, size_t bytes, void **out) {
    ArenaSlot *slot = nullptr;
    for (int i = 0; i < ctx->arena_n; i++)
        if (std::strncmp(ctx->arena_slot[i].name, name, sizeof(slot->name)) == 0) {
            slot = &ctx->arena_slot[i];
            break;
        }
    if (slot && slot->cap >= bytes) {
        *out = ctx->scratch_dev + slot->off;
        return VOX_OK;
    }
    if (!slot) {
        if (ctx->arena_n == kArenaSlots)
            return fail(ctx, VOX_ERR_UNSUPPORTED, "no-allocation call: too many scratch regions");
        slot = &ctx->arena_slot[ctx->arena_n++];
        std::strncpy(slot->name, name, sizeof(slot->name) - 1);
        slot->name[sizeof(slot->name) - 1] = '\0';
    }
    const size_t off = (ctx->arena_used + 255) & ~(size_t)255;
    if (off > kNoAllocScratchBytes || bytes > kNoAllocScratchBytes - off)
        return fail(ctx, VOX_ERR_UNSUPPORTED,
                    "no-allocation call: this grid needs more scratch than the " +
                        std::to_string(kNoAllocScratchBytes >> 20) +
                        " MiB the library holds per device (use vox_voxelize_nodes)");
    slot->off = off;
    slot->cap = bytes;
    ctx->arena_used = off + bytes;
    *out = ctx->scratch_dev + off;
    return VOX_OK;
}

template <class T>
vox_status buf(vox_ctx *ctx, const char *name, size_t count, T **out) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (ctx->arena) {
        void *p = nullptr;
        VOX_TRY(arena_buf(ctx, name, bytes, &p));
        *out = (T *)p;
        return VOX_OK;
    }
    Buf &b = ctx->bufs[name];
    if (bytes > b.cap) {
        if (b.p) {  // either stream may still use the old buffer
            VOX_HIP(hipStreamSynchronize(ctx->stream));
            if (ctx->side && ctx->side != ctx->stream) VOX_HIP(hipStreamSynchronize(ctx->side));
            if (ctx->side && ctx->stream == ctx->side) VOX_HIP(hipStreamSynchronize(ctx->main_stream));
            VOX_HIP(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        size_t want = bytes + bytes / 4;
        want = (want + 255) & ~(size_t)255;
        VOX_HIP(hipMalloc(&b.p, want));
        b.cap = want;
    }
    *out = (T *)b.p;
    return VOX_OK;
}

inline uint32_t grid_for(size_t n, uint32_t threads, uint32_t cap = 16384) {
    size_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    return (uint32_t)std::min<size_t>(g, cap);
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Clears with a kernel (graph-capture safe) rather than hipMemsetAsync.
vox_status fill(vox_ctx *ctx, uint32_t *p, size_t n, uint32_t v) {
    if (n == 0) return VOX_OK;
    k_fill_u32<<<grid_for(n, 256, 4096), 256, 0, ctx->stream>>>(p, n, v);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// Several fills in one launch (each launch on the stream costs a few
// microseconds whatever its size).
struct FillSeg {
    uint32_t *p;
    size_t n;
    uint32_t v;
};
constexpr int kMaxFillSegs = 6;
struct FillSegs {
    uint32_t *p[kMaxFillSegs];
    size_t n[kMaxFillSegs];
    uint32_t v[kMaxFillSegs];
    int count;
};
__global__ void k_fill_many(FillSegs s) {
    for (int k = 0; k < s.count; k++)
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.n[k];
             i += (size_t)gridDim.x * blockDim.x)
            s.p[k][i] = s.v[k];
}
vox_status fill_many(vox_ctx *ctx, std::initializer_list<FillSeg> segs) {
    FillSegs s;
    s.count = 0;
    size_t most = 0;
    for (const FillSeg &f : segs) {
        if (f.n == 0 || !f.p) continue;
        if (s.count == kMaxFillSegs) return fail(ctx, VOX_ERR_INVALID, "too many fill segments");
        s.p[s.count] = f.p;
        s.n[s.count] = f.n;
        s.v[s.count] = f.v;
        s.count++;
        most = std::max(most, f.n);
    }
    if (s.count == 0) return VOX_OK;
    k_fill_many<<<grid_for(most, 256, 4096), 256, 0, ctx->stream>>>(s);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// Exclusive scan of in[0,n) into out[0,n+1]: out[n] is the total mod 2^32,
// out[n+1] is 1 if the true total reached 2^32 (see k_scan_down).
vox_status scan(vox_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n, int depth = 0) {
    const uint32_t nblocks = n / kScanTile + 1;
    char name[32];
    const char *tag = ctx->stream == ctx->side ? "b" : "";  // one set per stream
    std::snprintf(name, sizeof(name), "scan%s_part%d", tag, depth);
    uint32_t *part;
    VOX_TRY(buf(ctx, name, 2 * (size_t)nblocks + 1, &part));
    if (nblocks <= kScanDirectTiles) {
        if (nblocks > 1) {
            k_scan_reduce<<<nblocks, kScanThreads, 0, ctx->stream>>>(in, n, part);
            VOX_HIP(hipGetLastError());
        }
        k_scan_down<<<nblocks, kScanThreads, 0, ctx->stream>>>(in, out, n, part);
        VOX_HIP(hipGetLastError());
        return VOX_OK;
    }
    k_scan_blocks<<<nblocks, kScanThreads, 0, ctx->stream>>>(in, out, n, part);
    VOX_HIP(hipGetLastError());
    if (nblocks > 1) {
        std::snprintf(name, sizeof(name), "scan%s_psum%d", tag, depth);
        uint32_t *psum;
        VOX_TRY(buf(ctx, name, nblocks + 2, &psum));
        VOX_TRY(scan(ctx, part, psum, nblocks, depth + 1));
        k_scan_add<<<nblocks, kScanThreads, 0, ctx->stream>>>(out, n, part, psum);
        VOX_HIP(hipGetLastError());
    }
    return VOX_OK;
}

vox_status read_u32(vox_ctx *ctx, const uint32_t *dev, size_t n, uint32_t *host) {
    if (n * 4 > kPinnedBytes) return fail(ctx, VOX_ERR_INVALID, "read_u32: read exceeds staging");
    VOX_HIP(hipMemcpyAsync(ctx->pinned, dev, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, ctx->pinned, n * 4);
    return VOX_OK;
}

int ilog2(uint32_t v) { return 31 - __builtin_clz(v); }

BrickDims brick_dims(const vox_grid &g, const int bdim[3]) {
    BrickDims d;
    d.nbx = (g.alloc[0] + bdim[0] - 1) / bdim[0];
    d.nby = (g.alloc[1] + bdim[1] - 1) / bdim[1];
    d.lwx = (uint32_t)ilog2((uint32_t)bdim[0] / 32u);
    d.lby = (uint32_t)ilog2((uint32_t)bdim[1]);
    d.lbz = (uint32_t)ilog2((uint32_t)bdim[2]);
    return d;
}

// A brick shape the surface and node passes support: powers of two, x a
// multiple of 64, at most 32 rows (y*z), at most 32 words per z plane (one
// mask word per brick layer), at least 4 planes (a layer then never spans
// two waves of the brick pass's epilogue), 256 to kMaxBrickVox voxels (the
// brick pass finishes 256 voxels per wave step), and (for the Cartesian node
// pass, whose threads own kNodeZ planes) z a multiple of kNodeZ.
bool brick_ok(const uint32_t b[3], bool cart) {
    for (int a = 0; a < 3; a++)
        if (b[a] == 0 || (b[a] & (b[a] - 1)) != 0) return false;
    const uint32_t nvox = b[0] * b[1] * b[2];
    // The brick pass's epilogue gives each wave whole layers: a layer of at
    // most 256 voxels tiles its 256-voxel steps; a larger one needs bricks at
    // least 4 deep (each wave then takes a multiple of a layer's steps).
    return b[0] >= 64 && b[1] * b[2] <= 32 && (b[0] / 32) * b[1] <= 32 &&
           (b[0] * b[1] <= 256 || b[2] >= 4) && nvox >= 256 && nvox <= (uint32_t)kMaxBrickVox &&
           (!cart || b[2] % kNodeZ == 0);
}

// ---------------------------------------------------------------------------
// Job tables.  Every call describes its grids as jobs (vox_kernels.h: Job):
// a Cartesian mesh is one job, an FCC lattice four (the shifted sub-grids).
// The per-pass item counts are summed into base arrays (njobs + 1 entries,
// the last one the total), so each stage is one launch over all jobs.
// ---------------------------------------------------------------------------
struct Plan {
    std::vector<Job> jobs;
    std::vector<SplitTable> splits;  // one per job
    std::vector<uint32_t> tri_b{0}, tile_b{0}, pad_b{0}, brick_b{0}, erode_b{0}, node_b{0};
    std::vector<uint32_t> fcc_first, fcc_b{0};
    std::vector<size_t> bits_off, surf_off, mat_off;  // into the call's arenas
    std::vector<uint32_t *> ext_bits;                  // caller's bit array or null
    size_t bits_words = 0, surf_words = 0, mat_bytes = 0;
    // PartialNode jobs: their LongNode arrays and winner-key grids
    std::vector<size_t> plong_off, key_off, pbits_off;
    size_t plong_bytes = 0, key_bytes = 0, pbits_words = 0;
    uint32_t maxW = 0;
    bool any_mats = false, any_partial = false;
};

// Device copies of a plan's tables.
struct PlanDev {
    Job *jobs;
    uint32_t *tri_b, *tile_b, *pad_b, *brick_b, *erode_b, *node_b, *fcc_first, *fcc_b;
    // A single job eroded in several subspaces: the pre-erosion copy of its
    // grid and the flags of the rows / planes where a later subspace starts.
    uint32_t *pre = nullptr;
    uint8_t *ystart = nullptr, *zstart = nullptr;
    // The solid stage runs per voxSplitRes y/z subspace (Job::sstart set).
    bool solid_split = false;
};

// Ext space and alloc layout of one grid: the whole plan, or one partition
// of it (a devConfig subspace or a slice; voxelizeWorker,
// src/host_side.cpp:901-1011).
struct Layout {
    uint32_t res[3];    // ext resolution (x, y, z)
    uint32_t alloc[3];  // allocated dims
    uint32_t oy, oz;    // alloc row / plane of ext row / plane 0
    double em[3];       // ext min vertex (before an FCC sub-grid shift)
    uint32_t padmask;   // Job::padmask
    bool zero_border, rot_ids;
};
Layout whole_layout(const vox_grid &g) {
    Layout L;
    for (int a = 0; a < 3; a++) {
        L.res[a] = g.res[a];
        L.alloc[a] = g.alloc[a];
        L.em[a] = g.min_vertex[a];
    }
    L.oy = L.oz = 1;
    L.padmask = 15u;
    L.zero_border = L.rot_ids = false;
    return L;
}
// The partition [pmin, pmax) of the plan's ext y/z space: its neighbour flags
// (a neighbour wherever the partition stops short of the grid's edge,
// 908-950), the ext partition reaching one row into each neighbour
// (994-1011), alloc dims partition + 2 and the ext min corner
// minVertex + extPartition.min * d.
Layout part_layout(const vox_grid &g, const uint32_t pmin[2], const uint32_t pmax[2]) {
    bool lo[2] = {false, false}, hi[2] = {false, false};
    for (int a = 0; a < 2; a++) {
        const uint32_t r = g.res[1 + a];
        if (pmin[a] == 0 && pmax[a] != r) hi[a] = true;
        else if (pmax[a] == r && pmin[a] != 0) lo[a] = true;
        else if (pmin[a] != 0 && pmax[a] != r) lo[a] = hi[a] = true;
    }
    Layout L;
    const uint32_t emin[2] = {pmin[0] - (lo[0] ? 1u : 0u), pmin[1] - (lo[1] ? 1u : 0u)};
    const uint32_t emax[2] = {pmax[0] + (hi[0] ? 1u : 0u), pmax[1] + (hi[1] ? 1u : 0u)};
    L.res[0] = g.res[0];
    L.res[1] = emax[0] - emin[0];
    L.res[2] = emax[1] - emin[1];
    L.alloc[0] = g.alloc[0];
    L.alloc[1] = pmax[0] - pmin[0] + 2;
    L.alloc[2] = pmax[1] - pmin[1] + 2;
    L.oy = lo[0] ? 0u : 1u;
    L.oz = lo[1] ? 0u : 1u;
    L.em[0] = g.min_vertex[0] + double(0u) * g.voxel_length;
    L.em[1] = g.min_vertex[1] + double(emin[0]) * g.voxel_length;
    L.em[2] = g.min_vertex[2] + double(emin[1]) * g.voxel_length;
    L.padmask = (lo[0] ? 0u : 1u) | (hi[0] ? 0u : 2u) | (lo[1] ? 0u : 4u) | (hi[1] ? 0u : 8u);
    L.zero_border = lo[0] || hi[0] || lo[1] || hi[1];
    L.rot_ids = false;
    return L;
}

// matSplitRes subspaces over extPartition (src/host_side.cpp:1107-1111).
vox_status split_table(vox_ctx *ctx, const uint32_t res[3], const uint32_t matSplit[3],
                       SplitTable &st) {
    std::memset(&st, 0, sizeof(st));
    for (int a = 0; a < 3; a++) {
        std::vector<Span> sp = split_axis_max(res[a], matSplit[a]);
        if (sp.empty() || sp.size() > (size_t)kMaxSplits)
            return fail(ctx, VOX_ERR_UNSUPPORTED, "matSplitRes gives too many subspaces");
        st.n[a] = (uint32_t)sp.size();
        for (size_t k = 0; k < sp.size(); k++) {
            st.lo[a][k] = sp[k].min;
            st.hi[a][k] = sp[k].max;
        }
    }
    return VOX_OK;
}

// The FCC interleave and the PartialNode expansion read surface bit words;
// the Cartesian node pass reads only the per-brick word masks.
inline bool job_needs_surf(const Job &J) { return J.pnodes != nullptr || J.kind >= 3; }

// Adds one grid with layout L; shift is the FCC sub-grid offset (zero for
// Cartesian).  cart: the job gets erosion and Cartesian node blocks.
vox_status add_job(vox_ctx *ctx, Plan &P, const float *V, const uint32_t *I, uint32_t nt,
                   const uint8_t *mats, const vox_grid &g, const Layout &L, const double shift[3],
                   bool use_mats, const uint32_t ms[3], const int bdim[3], bool cart, void *nodes,
                   int kind, uint32_t *ext_bits) {
    const uint32_t X = g.alloc[0], W = X / 32;
    if (W == 0) return fail(ctx, VOX_ERR_UNSUPPORTED, "empty x resolution");
    if (g.x_splits == 0 || X / g.x_splits > 1024)
        return fail(ctx, VOX_ERR_UNSUPPORTED,
                    "voxSplitRes.x gives x splits wider than 1024 voxels, where the reference's "
                    "word bitmask is undefined");
    if ((uint64_t)P.tri_b.back() + nt + 256 >= (1ull << 30))
        return fail(ctx, VOX_ERR_UNSUPPORTED, "more than 2^30 triangles in one call");
    Job J;
    std::memset(&J, 0, sizeof(J));
    J.V = V;
    J.I = I;
    J.mats = mats;
    J.nt = nt;
    J.tri_base = P.tri_b.back();
    P.tri_b.push_back(J.tri_base + (nt + 255u) / 256u * 256u);
    J.gd = GridDims{X, L.alloc[1], L.alloc[2], W};
    for (int a = 0; a < 3; a++) {
        J.res[a] = (int32_t)L.res[a];
        J.em[a] = L.em[a] + shift[a];
    }
    J.oy = L.oy;
    J.oz = L.oz;
    J.padmask = L.padmask;
    J.zero_border = L.zero_border ? 1 : 0;
    J.rot_ids = L.rot_ids ? 1 : 0;
    J.d = g.voxel_length;
    J.tilesY = (L.res[1] + 3) / 4;
    J.tilesZ = (L.res[2] + 3) / 4;
    J.tile_base = P.tile_b.back();
    P.tile_b.push_back(J.tile_base + J.tilesY * J.tilesZ);
    J.pad_base = P.pad_b.back();
    P.pad_b.push_back(J.pad_base + (pad_words(W, L.alloc[1], L.alloc[2]) + 255u) / 256u);
    J.m = F3{(float)J.em[0], (float)J.em[1], (float)J.em[2]};
    J.fd = (float)g.voxel_length;
    uint32_t nbt = 1;
    for (int a = 0; a < 3; a++) {
        J.bdim[a] = bdim[a];
        J.nb[a] = (L.alloc[a] + bdim[a] - 1) / bdim[a];
        nbt *= J.nb[a];
    }
    J.use_mats = use_mats ? 1 : 0;
    J.embed = nt < ctx->embed_limit ? 1 : 0;
    J.brick_base = P.brick_b.back();
    P.brick_b.push_back(J.brick_base + (use_mats ? nbt : 0u));
    vox_grid ga = g;
    for (int a = 0; a < 3; a++) ga.alloc[a] = L.alloc[a];
    J.bdn = brick_dims(ga, bdim);
    SplitTable st;
    std::memset(&st, 0, sizeof(st));
    if (use_mats) VOX_TRY(split_table(ctx, L.res, ms, st));  // only the surface pass reads it
    P.splits.push_back(st);
    J.erode_base = P.erode_b.back();
    J.node_base = P.node_b.back();
    P.erode_b.push_back(J.erode_base + (cart ? erode_blocks(J.gd) : 0u));
    P.node_b.push_back(J.node_base + (cart ? node_blocks(J.gd) : 0u));
    // PartialNode (kind 2) runs as LongNode into a LongNode array of the
    // call and is expanded into the caller's array after the node pass.
    const bool partial = kind == 2;
    J.nodes = partial ? nullptr : nodes;
    J.kind = partial ? 1 : kind;
    J.pnodes = partial ? nodes : nullptr;
    J.keygrid = nullptr;
    J.pbits = nullptr;
    P.plong_off.push_back(P.plong_bytes);
    P.key_off.push_back(P.key_bytes);
    P.pbits_off.push_back(P.pbits_words);
    if (partial) {
        P.plong_bytes += ((size_t)X * L.alloc[1] * L.alloc[2] * 2u + 255u) & ~(size_t)255u;
        P.pbits_words += ((size_t)W * L.alloc[1] * L.alloc[2] + 63u) & ~(size_t)63u;
        if (use_mats) P.key_bytes += (size_t)nbt * (size_t)(bdim[0] * bdim[1] * bdim[2]) * 4u;
        P.any_partial = true;
    }
    const size_t words = (size_t)W * L.alloc[1] * L.alloc[2];
    P.ext_bits.push_back(ext_bits);
    P.bits_off.push_back(P.bits_words);
    if (!ext_bits) P.bits_words += words;
    P.surf_off.push_back(P.surf_words);
    P.mat_off.push_back(P.mat_bytes);
    if (use_mats) {
        // Surface bit words in the layout of the solid grid (for the FCC and
        // PartialNode passes); one material byte per brick voxel,
        // brick-major.
        if (job_needs_surf(J)) P.surf_words += words;
        P.mat_bytes += (size_t)nbt * (size_t)(bdim[0] * bdim[1] * bdim[2]);
        P.any_mats = true;
    }
    P.maxW = std::max(P.maxW, W);
    P.jobs.push_back(J);
    return VOX_OK;
}

// A Cartesian mesh (one job) or an FCC lattice (four jobs).
vox_status add_mesh(vox_ctx *ctx, Plan &P, const float *V, const uint32_t *I, uint32_t nt,
                    const uint8_t *mats, const vox_grid &g, vox_node_kind kind, bool use_mats,
                    const uint32_t ms[3], void *nodes) {
    const double zero[3] = {0.0, 0.0, 0.0};
    if (kind == VOX_NODE_SHORT || kind == VOX_NODE_LONG || kind == VOX_NODE_PARTIAL)
        return add_job(ctx, P, V, I, nt, mats, g, whole_layout(g), zero, use_mats, ms,
                       ctx->cart_bdim, true, nodes, (int)kind, nullptr);
    // fccWorker (src/host_side.cpp:1262-1698): four sub-grids shifted by
    // (0,0,0), (h,h,0), (0,h,h), (h,0,h), h = d/2 (1436-1463).
    const double h = g.voxel_length / 2.0;
    const double shifts[4][3] = {{0.0, 0.0, 0.0}, {h, h, 0.0}, {0.0, h, h}, {h, 0.0, h}};
    const uint32_t first = (uint32_t)P.jobs.size();
    for (int k = 0; k < 4; k++)
        VOX_TRY(add_job(ctx, P, V, I, nt, mats, g, whole_layout(g), shifts[k], use_mats, ms,
                        ctx->fcc_bdim, false,
                        k == 0 ? nodes : nullptr, kind == VOX_NODE_SHORT_FCC ? 3 : 4, nullptr));
    P.fcc_first.push_back(first);
    P.fcc_b.push_back(P.fcc_b.back() + fcc_node_blocks(P.jobs[first].gd));
    return VOX_OK;
}

// Host-to-device upload through the context's pinned staging area; waits
// for the previous upload to leave the area first.
vox_status upload(vox_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return VOX_OK;
    if (bytes > ctx->up_cap && ctx->arena)
        return fail(ctx, VOX_ERR_UNSUPPORTED,
                    "no-allocation call: job table larger than the staging area");
    if (bytes > ctx->up_cap) {
        if (ctx->up) {
            VOX_HIP(hipEventSynchronize(ctx->up_ev));
            VOX_HIP(hipHostFree(ctx->up));
            ctx->up = nullptr;
        }
        const size_t want = (bytes + bytes / 2 + 4095) & ~(size_t)4095;
        VOX_HIP(hipHostMalloc(&ctx->up, want, hipHostMallocDefault));
        ctx->up_cap = want;
    } else {
        VOX_HIP(hipEventSynchronize(ctx->up_ev));
    }
    std::memcpy(ctx->up, src, bytes);
    VOX_HIP(hipMemcpyAsync(dst, ctx->up, bytes, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipEventRecord(ctx->up_ev, ctx->stream));
    return VOX_OK;
}

// Arenas of the call and the device job table (one upload: jobs, split
// tables, base arrays, erosion subspace flags).
vox_status plan_upload(vox_ctx *ctx, Plan &P, PlanDev &D, uint32_t splitY, uint32_t splitZ) {
    const uint32_t nj = (uint32_t)P.jobs.size(), nf = (uint32_t)P.fcc_first.size();
    uint32_t *bits = nullptr, *surf = nullptr, *bhist = nullptr, *keygrid = nullptr,
             *pbits = nullptr, *wmask = nullptr;
    uint8_t *matgrid = nullptr, *plong = nullptr;
    if (P.bits_words) VOX_TRY(buf(ctx, "a_bits", P.bits_words, &bits));
    if (P.plong_bytes) VOX_TRY(buf(ctx, "a_plong", P.plong_bytes, &plong));
    if (P.key_bytes) VOX_TRY(buf(ctx, "a_keygrid", P.key_bytes / 4u, &keygrid));
    if (P.pbits_words) VOX_TRY(buf(ctx, "a_pbits", P.pbits_words, &pbits));
    if (P.any_mats) {
        if (P.surf_words) VOX_TRY(buf(ctx, "a_surf", P.surf_words, &surf));
        VOX_TRY(buf(ctx, "a_matgrid", P.mat_bytes, &matgrid));
        VOX_TRY(buf(ctx, "f_brick_hist", P.brick_b.back(), &bhist));
        VOX_TRY(buf(ctx, "f_wmask", (size_t)P.brick_b.back() * 4u, &wmask));
    }
    std::vector<uint32_t> bases;
    for (auto *v : {&P.tri_b, &P.tile_b, &P.pad_b, &P.brick_b, &P.erode_b, &P.node_b})
        bases.insert(bases.end(), v->begin(), v->end());
    bases.insert(bases.end(), P.fcc_first.begin(), P.fcc_first.end());
    bases.insert(bases.end(), P.fcc_b.begin(), P.fcc_b.end());
    // A single job eroded in several subspaces: the ids of nodes next to a
    // later subspace read the pre-erosion grid (see k_nodes_cart); flags of
    // the rows and planes where a later subspace starts.
    std::vector<uint8_t> starts;
    D.pre = nullptr;
    D.ystart = D.zstart = nullptr;
    if (nj == 1 && P.erode_b.back() > 0) {
        const Job &J = P.jobs[0];
        std::vector<Span> eys = split_axis_max(J.gd.Y, splitY);
        std::vector<Span> ezs = split_axis_max(J.gd.Z, splitZ);
        if (eys.size() > 1 || ezs.size() > 1) {
            VOX_TRY(buf(ctx, "n_bits_pre", (size_t)J.gd.W * J.gd.Y * J.gd.Z, &D.pre));
            starts.assign(J.gd.Y + J.gd.Z, 0);
            for (size_t k = 1; k < eys.size(); k++) starts[eys[k].min] = 1;
            for (size_t k = 1; k < ezs.size(); k++) starts[J.gd.Y + ezs[k].min] = 1;
        }
    }
    // A single mesh (one Cartesian job, or the four sub-grids of one FCC
    // lattice, which share the ext resolution) whose voxSplitRes y/z gives
    // several solid subspaces: flags of the ext rows and planes where a later
    // subspace starts (Job::sstart).
    const size_t erode_flags = starts.size();
    const bool single_mesh = nj == 1 || (nj == 4 && nf == 1);
    if (single_mesh && nj > 0) {
        const Job &J = P.jobs[0];
        std::vector<Span> sys = split_axis_max((uint32_t)J.res[1], splitY);
        std::vector<Span> szs = split_axis_max((uint32_t)J.res[2], splitZ);
        if (sys.size() > 1 || szs.size() > 1) {
            const size_t o = starts.size();
            starts.resize(o + (size_t)J.res[1] + (size_t)J.res[2], 0);
            for (size_t k = 1; k < sys.size(); k++) starts[o + sys[k].min] = 1;
            for (size_t k = 1; k < szs.size(); k++) starts[o + (size_t)J.res[1] + szs[k].min] = 1;
        }
    }
    const bool solid_split = starts.size() > erode_flags;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t jb = al(sizeof(Job) * nj), sb = al(sizeof(SplitTable) * nj);
    const size_t bb = al(bases.size() * 4);
    uint8_t *dev;
    VOX_TRY(buf(ctx, "j_table", jb + sb + bb + starts.size(), &dev));
    const SplitTable *dsplit = (const SplitTable *)(dev + jb);
    for (uint32_t k = 0; k < nj; k++) {
        Job &J = P.jobs[k];
        J.bits = P.ext_bits[k] ? P.ext_bits[k] : bits + P.bits_off[k];
        J.splits = dsplit + k;
        J.sstart = solid_split ? dev + jb + sb + bb + erode_flags : nullptr;
        if (J.use_mats) {
            J.surf = job_needs_surf(J) ? surf + P.surf_off[k] : nullptr;
            J.bhist = bhist + J.brick_base;
            J.wmask = wmask + (size_t)J.brick_base * 4u;
            J.matgrid = matgrid + P.mat_off[k];
        }
        if (J.pnodes) {
            J.nodes = plong + P.plong_off[k];
            J.pbits = pbits + P.pbits_off[k];
            if (J.use_mats) J.keygrid = keygrid + P.key_off[k] / 4u;
        }
    }
    // One staging copy: jobs, split tables, base arrays, subspace flags.
    std::vector<uint8_t> host(jb + sb + bb + starts.size());
    std::memcpy(host.data(), P.jobs.data(), sizeof(Job) * nj);
    std::memcpy(host.data() + jb, P.splits.data(), sizeof(SplitTable) * nj);
    std::memcpy(host.data() + jb + sb, bases.data(), bases.size() * 4);
    if (!starts.empty()) std::memcpy(host.data() + jb + sb + bb, starts.data(), starts.size());
    VOX_TRY(upload(ctx, dev, host.data(), host.size()));
    D.jobs = (Job *)dev;
    uint32_t *b = (uint32_t *)(dev + jb + sb);
    D.tri_b = b;
    D.tile_b = (b += nj + 1);
    D.pad_b = (b += nj + 1);
    D.brick_b = (b += nj + 1);
    D.erode_b = (b += nj + 1);
    D.node_b = (b += nj + 1);
    D.fcc_first = (b += nj + 1);
    D.fcc_b = b + nf;
    if (erode_flags) {
        D.ystart = dev + jb + sb + bb;
        D.zstart = D.ystart + P.jobs[0].gd.Y;
    }
    D.solid_split = solid_split;
    return VOX_OK;
}

void record(vox_ctx *ctx, int i) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

// ---------------------------------------------------------------------------
// Solid stage over all jobs: per-triangle records, candidate tiles, counting
// sort by tile, fill.  Every grid is filled in one pass; with several
// voxSplitRes y/z subspaces each (triangle, tile) pair carries the tile rows
// whose subspace queues it, as the reference's per-subspace queues decide.
// ---------------------------------------------------------------------------
// Triangle records (vertices, class word, material) and slot jobs, read by
// both the solid and the surface stage; with fixed sizes the overflow flag
// is cleared here, before either stage can raise it.
vox_status records_stage(vox_ctx *ctx, const Plan &P, const PlanDev &D) {
    const uint32_t nj = (uint32_t)P.jobs.size();
    const uint32_t nslots = P.tri_b.back();
    VOX_TRY(buf(ctx, "tri_rec", (size_t)std::max<uint32_t>(nslots, 1) * 3, &ctx->rec));
    VOX_TRY(buf(ctx, "tri_job", nslots, &ctx->tri_job));
    if (ctx->fixed) VOX_TRY(fill(ctx, ctx->ovf, 1, 0u));
    if (nslots > 0) {
        k_tri_records<<<grid_for(nslots, 256), 256, 0, ctx->stream>>>(D.jobs, D.tri_b, nj, nslots,
                                                                       ctx->rec, ctx->tri_job);
        VOX_HIP(hipGetLastError());
    }
    return VOX_OK;
}

vox_status solid_stage(vox_ctx *ctx, const Plan &P, const PlanDev &D) {
    const uint32_t nj = (uint32_t)P.jobs.size();
    const uint32_t nslots = P.tri_b.back(), nTiles = P.tile_b.back();
    uint32_t *tri_job = ctx->tri_job, *cand_count, *cand_off, *tile_hist, *tile_off, *tile_cursor,
             *list;
    uint2 *pairs;
    uint4 *cbox;
    VOX_TRY(buf(ctx, "s_cand_count", nslots, &cand_count));
    VOX_TRY(buf(ctx, "s_cand_off", nslots + 2, &cand_off));
    VOX_TRY(buf(ctx, "s_cbox", nslots, &cbox));
    VOX_TRY(buf(ctx, "s_tile_hist", nTiles, &tile_hist));
    VOX_TRY(buf(ctx, "s_tile_off", nTiles + 2, &tile_off));
    VOX_TRY(buf(ctx, "s_tile_cursor", nTiles, &tile_cursor));
    if (nslots > 0) {
        k_solid_setup<<<grid_for(nslots, 256), 256, 0, ctx->stream>>>(D.jobs, D.tri_b, nj, nslots,
                                                                       ctx->rec, cand_count, cbox);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(scan(ctx, cand_count, cand_off, nslots));
    uint32_t total = 0;
    if (ctx->fixed) {
        total = ctx->cap_solid;
    } else {
        uint32_t h[2];
        VOX_TRY(read_u32(ctx, cand_off + nslots, 2, h));
        if (h[1])
            return fail(ctx, VOX_ERR_UNSUPPORTED,
                        "2^32 or more (triangle, tile) candidates in one call");
        total = h[0];
        ctx->last_solid = total;
    }
    uint32_t *ovf = ctx->fixed ? ctx->ovf : nullptr;
    uint32_t *first;
    uint16_t *pmask = nullptr, *list_mask = nullptr;
    VOX_TRY(buf(ctx, "s_pairs", total, &pairs));
    VOX_TRY(buf(ctx, "s_list", total, &list));
    if (D.solid_split) {
        VOX_TRY(buf(ctx, "s_pmask", total, &pmask));
        VOX_TRY(buf(ctx, "s_list_mask", total, &list_mask));
    }
    VOX_TRY(buf(ctx, "s_first", block_starts_size(total), &first));
    if (total > 0 && nslots > 0) {
        k_block_starts<<<grid_for(nslots, 256), 256, 0, ctx->stream>>>(cand_off, nslots, total,
                                                                        first);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(fill_many(ctx, {{tile_hist, nTiles, 0u}, {tile_cursor, nTiles, 0u}}));
    if (total > 0) {
        if (nj > 1)
            k_solid_candidates<true><<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                D.jobs, tri_job, nslots, ctx->rec, cand_off, cand_off + nslots, first, cbox,
                pairs, pmask, tile_hist, total, ovf);
        else
            k_solid_candidates<false><<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                D.jobs, tri_job, nslots, ctx->rec, cand_off, cand_off + nslots, first, cbox,
                pairs, pmask, tile_hist, total, ovf);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(scan(ctx, tile_hist, tile_off, nTiles));
    if (total > 0) {
        k_solid_scatter<<<grid_for(total, 256), 256, 0, ctx->stream>>>(
            pairs, pmask, cand_off + nslots, tile_off, tile_cursor, list, list_mask, total);
        VOX_HIP(hipGetLastError());
    }
    if (nTiles > 0) {
        k_solid_fill<<<nTiles, kWave, 16 * std::min(P.maxW, kFillSegWords) * 4, ctx->stream>>>(
            D.jobs, D.tile_b, nj, ctx->rec, tile_off, tile_hist, list, list_mask);
        VOX_HIP(hipGetLastError());
    }
    if (P.pad_b.back() > 0) {
        k_zero_pad_rows<<<P.pad_b.back(), 256, 0, ctx->stream>>>(D.jobs, D.pad_b, nj);
        VOX_HIP(hipGetLastError());
    }
    ctx->times.solid_pairs += total;  // candidate (triangle, tile) pairs
    return VOX_OK;
}

// Appends (features, measured cycles) of the listed bricks to the
// context's training samples.
vox_status record_samples(vox_ctx *ctx, const uint32_t *list, const uint32_t *n_list,
                          const float *feats, const uint64_t *cycles, uint32_t nbt) {
    uint32_t n = 0;
    VOX_TRY(read_u32(ctx, n_list, 1, &n));
    if (n == 0) return VOX_OK;
    std::vector<uint32_t> ids(n);
    std::vector<float> f((size_t)n * kCostFeatures);
    std::vector<uint64_t> cyc(nbt);
    VOX_HIP(hipMemcpyAsync(ids.data(), list, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    VOX_HIP(hipMemcpyAsync(f.data(), feats, f.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    VOX_HIP(hipMemcpyAsync(cyc.data(), cycles, (size_t)nbt * 8, hipMemcpyDeviceToHost,
                           ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->rec_feats.insert(ctx->rec_feats.end(), f.begin(), f.end());
    for (uint32_t q = 0; q < n; q++) ctx->rec_cycles.push_back((float)cyc[ids[q]]);
    return VOX_OK;
}

// ---------------------------------------------------------------------------
// Surface stage over the jobs that want materials: triangle classes and
// brick boxes, counting sort of (triangle, brick) pairs by brick, the brick
// order (plain or learned), the persistent brick pass, which leaves the
// surface words and the material grid the node passes read.
// ---------------------------------------------------------------------------
// Triangle slots from which the surface preparation runs on its own stream.
constexpr uint32_t kForkMinSlots = 16384;

// What the brick pass needs from the surface stage's preparation.
struct SurfPass {
    uint32_t *brick_off = nullptr, *brick_hist = nullptr, *blist = nullptr, *stats = nullptr;
    uint32_t *blist_order = nullptr, *next = nullptr;
    const uint32_t *n_list = nullptr, *hlist = nullptr, *hsub = nullptr, *n_heavy = nullptr;
    const uint8_t *skip = nullptr;
    float *feats = nullptr;
};

// Surface stage up to the brick pass: triangle setup, (triangle, brick)
// pairs binned by brick, the brick list and, for the learned schedule, its
// heavy bricks.  It only reads the triangle records, so it can run beside
// the solid stage.
vox_status surface_prep(vox_ctx *ctx, const Plan &P, const PlanDev &D, SurfPass &S) {
    const uint32_t nj = (uint32_t)P.jobs.size();
    const uint32_t nslots = P.tri_b.back(), nbt = P.brick_b.back();
    uint4 *tri_info, *tri_ext = nullptr;
    uint32_t *cand, *cand_off, *brick_hist, *brick_off, *brick_cursor, *blist, *stats;
    // The model is consulted only when there are more bricks than resident
    // brick-pass workgroups: with fewer, every brick starts at once and the
    // order cannot matter.
    const bool learned = ctx->schedule == VOX_SCHEDULE_LEARNED && ctx->cost_ready &&
                         nbt > (uint32_t)(ctx->n_cu * ctx->surf_blocks_per_cu);
    const bool features = learned || ctx->record;
    if (ctx->timing) ctx->times.learned_schedule = learned ? 1u : 0u;
    VOX_TRY(buf(ctx, "f_tri_info", nslots, &tri_info));
    if (features) VOX_TRY(buf(ctx, "f_tri_ext", nslots, &tri_ext));
    float4 *tri_rec = ctx->rec;
    VOX_TRY(buf(ctx, "f_cand", nslots, &cand));
    VOX_TRY(buf(ctx, "f_cand_off", nslots + 2, &cand_off));
    VOX_TRY(buf(ctx, "f_brick_hist", nbt, &brick_hist));
    VOX_TRY(buf(ctx, "f_brick_off", nbt + 2, &brick_off));
    VOX_TRY(buf(ctx, "f_brick_cursor", nbt, &brick_cursor));
    VOX_TRY(buf(ctx, "f_stats", 1, &stats));
    if (nslots > 0) {
        k_surf_setup<<<grid_for(nslots, 256), 256, 0, ctx->stream>>>(D.jobs, D.tri_b, nj, nslots,
                                                                      tri_info, cand, tri_ext,
                                                                      tri_rec);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(scan(ctx, cand, cand_off, nslots));
    uint32_t total = 0;
    if (ctx->fixed) {
        total = ctx->cap_surf;
    } else {
        uint32_t h[2];
        VOX_TRY(read_u32(ctx, cand_off + nslots, 2, h));
        if (h[1])
            return fail(ctx, VOX_ERR_UNSUPPORTED,
                        "2^32 or more (triangle, brick) candidates in one call");
        total = h[0];
        ctx->last_surf = total;
    }
    uint32_t *ovf = ctx->fixed ? ctx->ovf : nullptr;
    uint32_t *first;
    VOX_TRY(buf(ctx, "f_blist", total, &blist));
    VOX_TRY(buf(ctx, "f_first", block_starts_size(total), &first));
    if (total > 0 && nslots > 0) {
        k_block_starts<<<grid_for(nslots, 256), 256, 0, ctx->stream>>>(cand_off, nslots, total,
                                                                        first);
        VOX_HIP(hipGetLastError());
    }
    // Everything this stage and the brick pass count into starts at zero.
    uint32_t *wmask, *next, *cnt = nullptr;
    VOX_TRY(buf(ctx, "f_wmask", (size_t)nbt * 4u, &wmask));
    VOX_TRY(buf(ctx, "f_next", 1, &next));
    if (learned && !ctx->record) VOX_TRY(buf(ctx, "f_cnt", 2, &cnt));
    VOX_TRY(fill_many(ctx, {{brick_hist, nbt, 0u},
                            {brick_cursor, nbt, 0u},
                            {wmask, (size_t)nbt * 4u, 0u},
                            {stats, 1, 0u},
                            {next, 1, 0u},
                            {cnt, cnt ? 2u : 0u, 0u}}));
    auto bin = [&](int scatter) -> vox_status {
        if (total == 0) return VOX_OK;
        if (nj > 1)
            k_surf_bin<true><<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                D.jobs, ctx->tri_job, nslots, cand_off, cand_off + nslots, first, tri_info,
                brick_hist, brick_off, brick_cursor, blist, scatter, total, ovf);
        else
            k_surf_bin<false><<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                D.jobs, ctx->tri_job, nslots, cand_off, cand_off + nslots, first, tri_info,
                brick_hist, brick_off, brick_cursor, blist, scatter, total, ovf);
        VOX_HIP(hipGetLastError());
        return VOX_OK;
    };
    VOX_TRY(bin(0));
    VOX_TRY(scan(ctx, brick_hist, brick_off, nbt));
    // Non-empty bricks in index order (the plain schedule of the brick pass).
    uint32_t *flags, *blist_order, *blist_order_off;
    VOX_TRY(buf(ctx, "f_flags", nbt, &flags));
    VOX_TRY(buf(ctx, "f_order", nbt, &blist_order));
    VOX_TRY(buf(ctx, "f_order_off", nbt + 2, &blist_order_off));
    k_nonzero_flags<<<grid_for(nbt, 256), 256, 0, ctx->stream>>>(brick_hist, flags, nbt);
    VOX_HIP(hipGetLastError());
    VOX_TRY(scan(ctx, flags, blist_order_off, nbt));
    k_nonzero_list<<<grid_for(nbt, 256), 256, 0, ctx->stream>>>(brick_hist, blist_order_off,
                                                                blist_order, nbt);
    VOX_HIP(hipGetLastError());
    VOX_TRY(bin(1));
    if (ctx->timing) ctx->times.surface_pairs += total;
    // Learned schedule: the listed bricks screened by pair count, their cost
    // predicted by the MLP, the heavy ones on a list the brick pass takes
    // first (see vox_cost.hip).
    const uint32_t *n_list = blist_order_off + nbt;
    const uint32_t *hlist = nullptr, *hsubs = nullptr, *n_heavy = nullptr;
    const uint8_t *skip = nullptr;
    float *feats = nullptr;
    uint64_t *cycles = nullptr;
    if (ctx->record) VOX_TRY(buf(ctx, "f_feats", (size_t)nbt * kCostFeatures, &feats));
    if (ctx->record) {  // training samples: features of every listed brick
        k_brick_features<<<grid_for(nbt, 4, 65535), 256, 0, ctx->stream>>>(
            blist_order, nullptr, n_list, brick_hist, brick_off, blist, tri_info, tri_ext, D.jobs,
            D.brick_b, nj, feats);
        VOX_HIP(hipGetLastError());
    } else if (learned) {
        uint32_t *lcand, *heavy, *hsub;
        uint8_t *skipf;
        VOX_TRY(buf(ctx, "f_lcand", nbt, &lcand));
        VOX_TRY(buf(ctx, "f_heavy", (size_t)kMaxSplitParts * nbt, &heavy));
        VOX_TRY(buf(ctx, "f_hsub", (size_t)kMaxSplitParts * nbt, &hsub));
        VOX_TRY(buf(ctx, "f_skip", nbt, &skipf));
        const uint32_t g256 = grid_for(nbt, 256, 1024);
        k_cost_screen<<<g256, 256, 0, ctx->stream>>>(blist_order, n_list, brick_hist,
                                                     cand_off + nslots, ctx->screen_ratio, lcand,
                                                     cnt, skipf);
        VOX_HIP(hipGetLastError());
        // Cutting a brick repeats its triangle setup in every part; it pays
        // only where one long brick can end the pass last, i.e. when there
        // are few bricks per resident workgroup.
        const uint32_t resident = (uint32_t)(ctx->n_cu * ctx->surf_blocks_per_cu);
        const uint32_t per_wg = nbt / std::max(resident, 1u);
        const uint32_t max_parts =
            ctx->split_parts ? ctx->split_parts : per_wg >= 16u ? 1u : per_wg >= 4u ? 2u : 4u;
        k_cost_predict<<<grid_for(nbt, 4, (uint32_t)ctx->n_cu * 8u), 256, 0, ctx->stream>>>(
            ctx->cost_params, blist_order, lcand, cnt, ctx->heavy_cost, max_parts, brick_hist,
            brick_off, blist, tri_info, tri_ext, D.jobs, D.brick_b, nj, heavy, hsub, skipf);
        VOX_HIP(hipGetLastError());
        hlist = heavy;
        hsubs = hsub;
        n_heavy = cnt + 1;
        skip = skipf;
    }
    (void)cycles;
    S.brick_off = brick_off;
    S.brick_hist = brick_hist;
    S.blist = blist;
    S.stats = stats;
    S.blist_order = blist_order;
    S.next = next;
    S.n_list = n_list;
    S.hlist = hlist;
    S.hsub = hsubs;
    S.n_heavy = n_heavy;
    S.skip = skip;
    S.feats = feats;
    return VOX_OK;
}

// The brick pass (after the solid stage: it ORs surface bits into the solid
// grid).
vox_status surface_brick(vox_ctx *ctx, const Plan &P, const PlanDev &D, const SurfPass &S) {
    const uint32_t nj = (uint32_t)P.jobs.size(), nbt = P.brick_b.back();
    uint64_t *cycles = nullptr;
    if (ctx->record) VOX_TRY(buf(ctx, "f_cycles", nbt, &cycles));
    const uint32_t blocks = std::min<uint32_t>(nbt, (uint32_t)(ctx->n_cu * ctx->surf_blocks_per_cu));
    if (P.any_partial)
        k_surf_brick<true><<<blocks, 256, 0, ctx->stream>>>(
            D.jobs, D.brick_b, nj, ctx->rec, S.brick_off, S.brick_hist, S.blist,
            ctx->timing ? S.stats : nullptr, S.blist_order, S.n_list, S.hlist, S.hsub, S.n_heavy,
            S.skip, S.next, cycles);
    else
        k_surf_brick<false><<<blocks, 256, 0, ctx->stream>>>(
            D.jobs, D.brick_b, nj, ctx->rec, S.brick_off, S.brick_hist, S.blist,
            ctx->timing ? S.stats : nullptr, S.blist_order, S.n_list, S.hlist, S.hsub, S.n_heavy,
            S.skip, S.next, cycles);
    VOX_HIP(hipGetLastError());
    if (ctx->timing) {
        uint32_t vox = 0;
        VOX_TRY(read_u32(ctx, S.stats, 1, &vox));
        ctx->times.surface_voxels += vox;
    }
    if (ctx->record)
        VOX_TRY(record_samples(ctx, S.blist_order, S.n_list, S.feats, cycles, nbt));
    return VOX_OK;
}

// Erosion to the fixed point over every Cartesian job, in place.  A single
// job may have several alloc y/z subspaces, eroded one after another in the
// reference's order; jobs of a batch have one each.  Passes are launched
// kBatch at a time; a job's pass after one that changed nothing in it exits
// at once, so one host read per batch of passes suffices.  With fixed sizes
// (graph capture) each subspace gets ctx->n_passes passes and no read-back;
// the last pass of every subspace raises the not-converged flag if it still
// changed something.
vox_status erode_stage(vox_ctx *ctx, const Plan &P, const PlanDev &D, uint32_t splitY,
                       uint32_t splitZ) {
    constexpr uint32_t kBatch = 4;
    const uint32_t nj = (uint32_t)P.jobs.size();
    const uint32_t nblocks = P.erode_b.back();
    ctx->last_changed = nullptr;
    if (nblocks == 0) return VOX_OK;
    uint32_t *changes, *totals, *stamp, *notconv = nullptr;
    const uint32_t nbuf = ctx->fixed ? ctx->n_passes : kBatch;  // passes per launch group
    VOX_TRY(buf(ctx, "e_changes", (size_t)nbuf * nj, &changes));
    VOX_TRY(buf(ctx, "e_totals", nbuf, &totals));
    VOX_TRY(buf(ctx, "e_stamp", nblocks, &stamp));
    if (ctx->fixed) {
        VOX_TRY(buf(ctx, "e_notconv", 1, &notconv));
        ctx->last_changed = notconv;
    }
    bool cleared = false;  // stamp and notconv go with the first group's flags
    std::vector<Span> ys{{0u, UINT32_MAX}}, zs{{0u, UINT32_MAX}};
    if (nj == 1) {
        ys = split_axis_max(P.jobs[0].gd.Y, splitY);
        zs = split_axis_max(P.jobs[0].gd.Z, splitZ);
    }
    // Persistent passes: at most the resident workgroups (a multiple of 8).
    const uint32_t grid = std::min<uint32_t>(
        xcd_grid(nblocks), xcd_grid((uint32_t)(ctx->n_cu * ctx->erode_blocks_per_cu)));
    uint32_t passes = 0, most = 0, gpass = 0;
    for (const Span &zsp : zs)
        for (const Span &ysp : ys) {
            bool first = true;
            uint32_t sub = 0;  // passes this subspace needed
            for (int guard = 0; guard < 1000000; guard++) {
                VOX_TRY(fill_many(ctx, {{changes, (size_t)nbuf * nj, 0u},
                                        {totals, nbuf, 0u},
                                        {cleared ? nullptr : stamp, nblocks, 0u},
                                        {cleared ? nullptr : notconv, 1, 0u}}));
                cleared = true;
                for (uint32_t k = 0; k < nbuf; k++) {
                    k_erode<<<grid, 256, 0, ctx->stream>>>(
                        D.jobs, D.erode_b, nj, ysp.min, ysp.max, zsp.min, zsp.max, stamp,
                        changes, totals, k == nbuf - 1 ? notconv : nullptr, k, gpass + k,
                        (first && k == 0) ? 1 : 0);
                    VOX_HIP(hipGetLastError());
                }
                gpass += nbuf;
                first = false;
                if (ctx->fixed) {
                    sub = nbuf;
                    break;
                }
                uint32_t h[kBatch];
                VOX_TRY(read_u32(ctx, totals, kBatch, h));
                uint32_t k0 = kBatch;
                for (uint32_t k = 0; k < kBatch; k++)
                    if (h[k] == 0u) {
                        k0 = k;
                        break;
                    }
                if (k0 < kBatch) {
                    sub += k0 + 1;
                    break;
                }
                sub += kBatch;
            }
            passes += sub;
            most = std::max(most, sub);
        }
    ctx->times.erosion_passes += passes;
    if (!ctx->fixed) ctx->last_passes = most;
    return VOX_OK;
}

// Node passes: Cartesian jobs and FCC lattices.  With several erosion
// subspaces (single job only) the ids next to a later subspace read the
// pre-erosion grid (see k_nodes_cart).
vox_status node_stage(vox_ctx *ctx, const Plan &P, const PlanDev &D, const uint32_t *pre,
                      const uint8_t *ystart, const uint8_t *zstart) {
    const uint32_t nj = (uint32_t)P.jobs.size(), nf = (uint32_t)P.fcc_first.size();
    if (P.node_b.back() > 0) {
        if (pre)
            k_nodes_cart_pre<<<xcd_grid(P.node_b.back()), 256, 0, ctx->stream>>>(
                D.jobs, D.node_b, nj, pre, ystart, zstart);
        else
            k_nodes_cart<<<xcd_grid(P.node_b.back()), 256, 0, ctx->stream>>>(D.jobs, D.node_b,
                                                                              nj);
        VOX_HIP(hipGetLastError());
    }
    if (nf > 0) {
        k_nodes_fcc<<<xcd_grid(P.fcc_b.back()), dim3(kWave, 4), 0, ctx->stream>>>(
            D.jobs, D.fcc_first, D.fcc_b, nf);
        VOX_HIP(hipGetLastError());
    }
    for (uint32_t j = 0; j < nj; j++) {
        const Job &J = P.jobs[j];
        if (!J.pnodes) continue;
        const size_t n = (size_t)J.gd.X * J.gd.Y * J.gd.Z;
        k_nodes_partial<<<grid_for(n, 256, 16384), 256, 0, ctx->stream>>>(D.jobs, j);
        VOX_HIP(hipGetLastError());
    }
    return VOX_OK;
}

// The whole pipeline for a plan.  splitY/splitZ: voxSplitRes y/z (erosion
// subspaces; single-job plans only).
vox_status run_stages(vox_ctx *ctx, Plan &P, const PlanDev &D, uint32_t splitY,
                      uint32_t splitZ);

vox_status run_plan(vox_ctx *ctx, Plan &P, uint32_t splitY, uint32_t splitZ) {
    PlanDev D;
    record(ctx, 0);
    VOX_TRY(plan_upload(ctx, P, D, splitY, splitZ));
    return run_stages(ctx, P, D, splitY, splitZ);
}

// The stages after the job-table upload (what a graph captures).
vox_status run_stages(vox_ctx *ctx, Plan &P, const PlanDev &D, uint32_t splitY,
                      uint32_t splitZ) {
    VOX_TRY(records_stage(ctx, P, D));
    // Fork: the surface preparation on the side stream, the solid stage on
    // the main one; they join before the brick pass.  (Inside a graph
    // capture the side stream joins the capture through the fork event.)
    SurfPass S;
    // Small meshes (a few thousand triangles) have too little of either
    // stage to overlap; the fork's extra dependencies would cost more.
    const bool fork = P.any_mats && P.tri_b.back() >= kForkMinSlots;
    if (fork) {
        VOX_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
        VOX_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        hipStream_t main = ctx->stream;
        ctx->stream = ctx->side;
        const vox_status st = surface_prep(ctx, P, D, S);
        ctx->stream = main;
        VOX_TRY(st);
        VOX_HIP(hipEventRecord(ctx->ev_join, ctx->side));
    }
    VOX_TRY(solid_stage(ctx, P, D));
    record(ctx, 1);
    if (fork) {
        VOX_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
        VOX_TRY(surface_brick(ctx, P, D, S));
    } else if (P.any_mats) {
        VOX_TRY(surface_prep(ctx, P, D, S));
        VOX_TRY(surface_brick(ctx, P, D, S));
    }
    record(ctx, 2);
    if (D.pre) {
        const Job &J = P.jobs[0];
        VOX_HIP(hipMemcpyAsync(D.pre, J.bits, (size_t)J.gd.W * J.gd.Y * J.gd.Z * 4,
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    // PartialNode keeps r = float(solid) of the nodes erosion removes: the
    // bits as they were before erosion.
    for (const Job &J : P.jobs)
        if (J.pbits)
            VOX_HIP(hipMemcpyAsync(J.pbits, J.bits, (size_t)J.gd.W * J.gd.Y * J.gd.Z * 4,
                                   hipMemcpyDeviceToDevice, ctx->stream));
    VOX_TRY(erode_stage(ctx, P, D, splitY, splitZ));
    record(ctx, 3);
    VOX_TRY(node_stage(ctx, P, D, D.pre, D.ystart, D.zstart));
    record(ctx, 4);
    return VOX_OK;
}

vox_status finish_times(vox_ctx *ctx) {
    if (!ctx->timing) return VOX_OK;
    VOX_HIP(hipEventSynchronize(ctx->ev[4]));
    (void)hipEventElapsedTime(&ctx->times.solid_fill_ms, ctx->ev[0], ctx->ev[1]);
    (void)hipEventElapsedTime(&ctx->times.surface_ms, ctx->ev[1], ctx->ev[2]);
    (void)hipEventElapsedTime(&ctx->times.erode_ms, ctx->ev[2], ctx->ev[3]);
    (void)hipEventElapsedTime(&ctx->times.nodes_ms, ctx->ev[3], ctx->ev[4]);
    (void)hipEventElapsedTime(&ctx->times.total_ms, ctx->ev[0], ctx->ev[4]);
    return VOX_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int vox_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *vox_build_info(void) { return "voxelizer_amd gfx950 (wave64, HIP)"; }

vox_status vox_create(int device, vox_ctx **out) {
    if (!out) return VOX_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return VOX_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return VOX_ERR_NO_DEVICE;
    vox_ctx *ctx = new vox_ctx();
    ctx->dev = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        (ctx->main_stream = ctx->stream, false) ||
        hipHostMalloc((void **)&ctx->pinned, kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return VOX_ERR_HIP;
    }
    for (auto &e : ctx->ev) (void)hipEventCreate(&e);
    if (hipEventCreateWithFlags(&ctx->up_ev, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreate(&ctx->side) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
        vox_destroy(ctx);
        return VOX_ERR_HIP;
    }
    if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->surf_blocks_per_cu, k_surf_brick<false>, 256,
                                                     0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->erode_blocks_per_cu, k_erode, 256, 0) !=
            hipSuccess) {
        vox_destroy(ctx);
        return VOX_ERR_HIP;
    }
    ctx->n_cu = std::max(ctx->n_cu, 1);
    // Set-up of the no-allocation entry point: this device's copy of the
    // scratch array, and a pinned staging area for job tables.
    {
        hipError_t e = hipSuccess;
        ctx->scratch_dev = (uint8_t *)scratch_on(device, &e);
        if (e != hipSuccess) {
            vox_destroy(ctx);
            return VOX_ERR_HIP;
        }
        constexpr size_t kStaging = 64 * 1024;
        if (hipHostMalloc(&ctx->up, kStaging, hipHostMallocDefault) != hipSuccess) {
            vox_destroy(ctx);
            return VOX_ERR_HIP;
        }
        ctx->up_cap = kStaging;
    }
    if (const char *e = std::getenv("VOX_EMBED_TRIS"))
        ctx->embed_limit = (uint32_t)std::min<unsigned long>(std::strtoul(e, nullptr, 10), 1ul << 22);
    ctx->surf_blocks_per_cu = std::max(ctx->surf_blocks_per_cu, 1);
    ctx->erode_blocks_per_cu = std::max(ctx->erode_blocks_per_cu, 1);
    // The node pass evaluates the boundary-id table as logic; check it.
    uint8_t lut[64] = {0};
    for (auto &pr : kOrientPairs) lut[pr[0]] = pr[1];
    for (uint32_t p = 0; p < 64; p++) {
        const IdPlanes o = orient_planes(p & 1u, (p >> 1) & 1u, (p >> 2) & 1u, (p >> 3) & 1u,
                                         (p >> 4) & 1u, (p >> 5) & 1u);
        uint32_t id = 0;
        for (int b = 0; b < 5; b++) id |= (o.b[b] & 1u) << b;
        if (id != lut[p]) {
            vox_destroy(ctx);
            return VOX_ERR_INVALID;
        }
    }
    *out = ctx;
    return VOX_OK;
}

void vox_destroy(vox_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->dev);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (auto &kv : ctx->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    if (ctx->own_mesh) {
        (void)hipFree((void *)ctx->V);
        (void)hipFree((void *)ctx->I);
    }
    if (ctx->own_mats) (void)hipFree((void *)ctx->mats);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (vox_ctx *w : ctx->workers) vox_destroy(w);
    ctx->workers.clear();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->up) (void)hipHostFree(ctx->up);
    if (ctx->up_ev) (void)hipEventDestroy(ctx->up_ev);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->cost_params) (void)hipFree(ctx->cost_params);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *vox_last_error(const vox_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

vox_status vox_sync(vox_ctx *ctx) {
    if (!ctx) return VOX_ERR_INVALID;
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    return VOX_OK;
}

vox_status vox_alloc(vox_ctx *ctx, size_t bytes, void **dev_out) {
    if (!ctx || !dev_out) return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    VOX_HIP(hipMalloc(dev_out, std::max<size_t>(bytes, 1)));
    return VOX_OK;
}
vox_status vox_free(vox_ctx *ctx, void *dev) {
    if (!ctx) return VOX_ERR_INVALID;
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    VOX_HIP(hipFree(dev));
    return VOX_OK;
}
vox_status vox_memcpy_h2d(vox_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return VOX_ERR_INVALID;
    VOX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    return VOX_OK;
}
vox_status vox_memcpy_d2h(vox_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return VOX_ERR_INVALID;
    VOX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    return VOX_OK;
}

static void drop_mesh(vox_ctx *ctx) {
    if (ctx->own_mesh) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree((void *)ctx->V);
        (void)hipFree((void *)ctx->I);
    }
    ctx->V = nullptr;
    ctx->I = nullptr;
    ctx->own_mesh = false;
}

vox_status vox_set_mesh(vox_ctx *ctx, const float *vertices, uint32_t nv, const uint32_t *indices,
                        uint32_t nt) {
    if (!ctx || (!vertices && nv) || (!indices && nt)) return VOX_ERR_INVALID;
    for (uint64_t i = 0; i < 3ull * nt; i++)
        if (indices[i] >= nv) return fail(ctx, VOX_ERR_INVALID, "vertex index out of range");
    drop_mesh(ctx);
    float *v;
    uint32_t *ix;
    VOX_HIP(hipMalloc((void **)&v, std::max<size_t>(12ull * nv, 4)));
    VOX_HIP(hipMalloc((void **)&ix, std::max<size_t>(12ull * nt, 4)));
    if (nv) VOX_HIP(hipMemcpyAsync(v, vertices, 12ull * nv, hipMemcpyHostToDevice, ctx->stream));
    if (nt) VOX_HIP(hipMemcpyAsync(ix, indices, 12ull * nt, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->V = v;
    ctx->I = ix;
    ctx->nv = nv;
    ctx->nt = nt;
    ctx->own_mesh = true;
    return VOX_OK;
}

vox_status vox_set_mesh_device(vox_ctx *ctx, const float *vertices_dev, uint32_t nv,
                               const uint32_t *indices_dev, uint32_t nt) {
    if (!ctx || (!vertices_dev && nv) || (!indices_dev && nt)) return VOX_ERR_INVALID;
    drop_mesh(ctx);
    ctx->V = vertices_dev;
    ctx->I = indices_dev;
    ctx->nv = nv;
    ctx->nt = nt;
    return VOX_OK;
}

vox_status vox_set_materials(vox_ctx *ctx, const uint8_t *materials, uint32_t n_unique) {
    if (!ctx) return VOX_ERR_INVALID;
    if (ctx->own_mats) {
        VOX_HIP(hipStreamSynchronize(ctx->stream));
        VOX_HIP(hipFree((void *)ctx->mats));
    }
    ctx->mats = nullptr;
    ctx->own_mats = false;
    ctx->n_unique = n_unique;
    if (!materials) return VOX_OK;
    uint8_t *m;
    VOX_HIP(hipMalloc((void **)&m, std::max<size_t>(ctx->nt, 1)));
    if (ctx->nt)
        VOX_HIP(hipMemcpyAsync(m, materials, ctx->nt, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->mats = m;
    ctx->own_mats = true;
    return VOX_OK;
}

vox_status vox_set_materials_device(vox_ctx *ctx, const uint8_t *materials_dev,
                                    uint32_t n_unique) {
    if (!ctx) return VOX_ERR_INVALID;
    if (ctx->own_mats) {
        VOX_HIP(hipStreamSynchronize(ctx->stream));
        VOX_HIP(hipFree((void *)ctx->mats));
    }
    ctx->mats = materials_dev;
    ctx->own_mats = false;
    ctx->n_unique = n_unique;
    return VOX_OK;
}

namespace {

// Dimensions in FP64 exactly as determineDimensions() /
// determineDimensions(double) (src/host_side.cpp:2497-2609) and
// splitResolutionByMaxDim + adjustResolution (src/host_side.cpp:730-738,
// 2620-2634) compute them from the bounding box.
void plan_from_bbox(const double mn[3], const double mx[3], vox_plan_mode mode, double param,
                    uint32_t vsx, vox_grid *g) {
    auto ceil_u = [](double v) -> uint32_t {
        double c = std::ceil(v);
        if (c != c || c <= -1.0) return 0u;
        if (c >= 4294967296.0) return UINT32_MAX;
        return (uint32_t)c;
    };
    double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    uint32_t rx, ry, rz;
    double d;
    if (mode == VOX_PLAN_MAX_DIMENSION) {
        uint32_t R = (uint32_t)param;
        if (dx > dy) {
            if (dx > dz) {
                d = dx / double(R - 1);
                ry = ceil_u(dy / d) + 1;
                rz = ceil_u(dz / d) + 1;
                rx = R + 2;
            } else {
                d = dz / double(R - 1);
                rz = R;
                rx = ceil_u(dx / d) + 1 + 2;
                ry = ceil_u(dy / d) + 1;
            }
        } else if (dy > dz) {
            d = dy / double(R - 1);
            ry = R;
            rx = ceil_u(dx / d) + 1 + 2;
            rz = ceil_u(dz / d) + 1;
        } else {
            d = dz / double(R - 1);
            rz = R;
            rx = ceil_u(dx / d) + 1 + 2;
            ry = ceil_u(dy / d) + 1;
        }
    } else {
        d = param;
        rx = ceil_u(dx / d) + 1 + 2;
        ry = ceil_u(dy / d) + 1;
        rz = ceil_u(dz / d) + 1;
    }
    uint32_t xs = (rx + vsx - 1) / vsx;
    uint32_t xmod = 32u * xs;
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
}

bool plan_args_ok(vox_plan_mode mode, double param, uint32_t vsx) {
    if (vsx == 0) return false;
    if (mode == VOX_PLAN_MAX_DIMENSION) return param >= 2.0 && param <= 4294967295.0;
    if (mode == VOX_PLAN_CUBE_LENGTH) return param > 0.0;
    return false;
}

}  // namespace

// Bounding box on the device (calculateBoundingBox, src/host_side.cpp:
// 2462-2488, as an ordered-key min/max reduction); dimensions on the host.
vox_status vox_plan_grid(vox_ctx *ctx, vox_plan_mode mode, double param, uint32_t vsx,
                         vox_grid *g) {
    if (!ctx || !g || vsx == 0) return VOX_ERR_INVALID;
    if (ctx->nv == 0) return fail(ctx, VOX_ERR_INVALID, "no vertices");
    if (!plan_args_ok(mode, param, vsx))
        return fail(ctx, VOX_ERR_INVALID,
                    mode == VOX_PLAN_MAX_DIMENSION ? "maxDimension must be >= 2"
                                                   : "cubeLength must be > 0");
    VOX_HIP(hipSetDevice(ctx->dev));
    uint32_t *bb;
    VOX_TRY(buf(ctx, "p_bbox", 6, &bb));
    VOX_TRY(fill_many(ctx, {{bb, 3, UINT32_MAX}, {bb + 3, 3, 0u}}));
    k_bbox<<<grid_for(ctx->nv, 256, 256), 256, 0, ctx->stream>>>(ctx->V, ctx->nv, bb);
    VOX_HIP(hipGetLastError());
    uint32_t h[6];
    VOX_TRY(read_u32(ctx, bb, 6, h));
    double mn[3], mx[3];
    for (int c = 0; c < 3; c++) {
        uint32_t a = h[c], b = h[3 + c];
        uint32_t ua = (a & 0x80000000u) ? (a & 0x7FFFFFFFu) : ~a;
        uint32_t ub = (b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b;
        float fa, fb;
        std::memcpy(&fa, &ua, 4);
        std::memcpy(&fb, &ub, 4);
        mn[c] = fa;
        mx[c] = fb;
    }
    plan_from_bbox(mn, mx, mode, param, vsx, g);
    return VOX_OK;
}

// Host bounding box with the reference's strict comparisons
// (src/host_side.cpp:2462-2488); no device is touched.
vox_status vox_plan_grid_host(const float *v, uint32_t nv, vox_plan_mode mode, double param,
                              uint32_t vsx, vox_grid *g) {
    if (!v || !g || nv == 0 || !plan_args_ok(mode, param, vsx)) return VOX_ERR_INVALID;
    double mn[3] = {v[0], v[1], v[2]}, mx[3] = {v[0], v[1], v[2]};
    for (uint32_t i = 1; i < nv; i++)
        for (int c = 0; c < 3; c++) {
            const double x = v[3 * (size_t)i + c];
            if (x > mx[c]) mx[c] = x;
            if (x < mn[c]) mn[c] = x;
        }
    plan_from_bbox(mn, mx, mode, param, vsx, g);
    return VOX_OK;
}

void vox_node_dims(const vox_grid *g, vox_node_kind kind, uint32_t dims[3]) {
    const bool fcc = kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
    dims[0] = g->alloc[0] * (fcc ? 2 : 1);
    dims[1] = g->alloc[1];
    dims[2] = g->alloc[2] * (fcc ? 2 : 1);
}
size_t vox_node_count(const vox_grid *g, vox_node_kind kind) {
    uint32_t d[3];
    vox_node_dims(g, kind, d);
    return (size_t)d[0] * d[1] * d[2];
}
size_t vox_node_size(vox_node_kind kind) {
    switch (kind) {
        case VOX_NODE_SHORT: return 1;
        case VOX_NODE_LONG: return 2;
        case VOX_NODE_PARTIAL: return 8;
        case VOX_NODE_SHORT_FCC: return 2;
        case VOX_NODE_LONG_FCC: return 4;
    }
    return 0;
}

static bool grid_ok(const vox_grid *g) {
    return g && g->alloc[0] % 32 == 0 && g->alloc[0] > 0 && g->alloc[1] >= 3 && g->alloc[2] >= 3 &&
           g->alloc[1] == g->res[1] + 2 && g->alloc[2] == g->res[2] + 2 &&
           g->alloc[0] == g->res[0] && g->voxel_length > 0.0 && g->x_splits > 0 &&
           (double)g->alloc[0] * g->alloc[1] * g->alloc[2] < 4294967296.0;
}

vox_status vox_solid_bits(vox_ctx *ctx, const vox_grid *g, const double shift[3],
                          const uint32_t split_yz[2], uint32_t *bits_dev) {
    if (!ctx || !grid_ok(g) || !bits_dev) return VOX_ERR_INVALID;
    if (!ctx->V && ctx->nv) return fail(ctx, VOX_ERR_INVALID, "no mesh");
    VOX_HIP(hipSetDevice(ctx->dev));
    ctx->times = vox_stage_times{};
    const double zero[3] = {0.0, 0.0, 0.0};
    const uint32_t sy = split_yz ? split_yz[0] : 512u, sz = split_yz ? split_yz[1] : 512u;
    if (sy == 0 || sz == 0) return VOX_ERR_INVALID;
    // Solid subspaces come from split_yz (over the ext resolution, as
    // performVoxelization's extYzSplits); the material split is unused here.
    const uint32_t ms[3] = {512u, 512u, 512u};
    Plan P;
    VOX_TRY(add_job(ctx, P, ctx->V, ctx->I, ctx->nt, nullptr, *g, whole_layout(*g),
                    shift ? shift : zero, false, ms,
                    ctx->cart_bdim, false, nullptr, 0, bits_dev));
    PlanDev D;
    record(ctx, 0);
    VOX_TRY(plan_upload(ctx, P, D, sy, sz));
    VOX_TRY(records_stage(ctx, P, D));
    VOX_TRY(solid_stage(ctx, P, D));
    record(ctx, 1);
    if (ctx->timing) {
        VOX_HIP(hipEventSynchronize(ctx->ev[1]));
        (void)hipEventElapsedTime(&ctx->times.solid_fill_ms, ctx->ev[0], ctx->ev[1]);
        ctx->times.total_ms = ctx->times.solid_fill_ms;
    }
    return VOX_OK;
}

static bool node_kind_ok(vox_node_kind kind) {
    return kind == VOX_NODE_SHORT || kind == VOX_NODE_LONG || kind == VOX_NODE_PARTIAL ||
           kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
}

static vox_status voxelize_nodes_impl(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                                      int use_materials, const uint32_t vox_split[3],
                                      const uint32_t mat_split[3], void *nodes_dev);

vox_status vox_voxelize_nodes(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                              int use_materials, const uint32_t vox_split[3],
                              const uint32_t mat_split[3], void *nodes_dev) {
    if (!ctx || !grid_ok(g) || !nodes_dev) return VOX_ERR_INVALID;
    return voxelize_nodes_impl(ctx, g, kind, use_materials, vox_split, mat_split, nodes_dev);
}

// The same pipeline with every buffer carved from this device's scratch
// array; the job table goes through the staging area vox_create allocated.
vox_status vox_voxelize_nodes_noalloc(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                                      int use_materials, const uint32_t vox_split[3],
                                      const uint32_t mat_split[3], void *nodes_dev) {
    if (!ctx || !grid_ok(g) || !nodes_dev) return VOX_ERR_INVALID;
    if (!ctx->scratch_dev)
        return fail(ctx, VOX_ERR_UNSUPPORTED, "no scratch array on this device");
    if (ctx->record)
        return fail(ctx, VOX_ERR_UNSUPPORTED,
                    "cost recording is not available in the no-allocation call");
    ctx->arena = true;
    ctx->arena_used = 0;
    ctx->arena_n = 0;
    const vox_status st =
        voxelize_nodes_impl(ctx, g, kind, use_materials, vox_split, mat_split, nodes_dev);
    ctx->arena = false;
    return st;
}

static vox_status voxelize_nodes_impl(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                                      int use_materials, const uint32_t vox_split[3],
                                      const uint32_t mat_split[3], void *nodes_dev) {
    if (!node_kind_ok(kind)) return VOX_ERR_INVALID;
    if (use_materials && !ctx->mats)
        return fail(ctx, VOX_ERR_INVALID, "materials requested but none set");
    VOX_HIP(hipSetDevice(ctx->dev));
    const uint32_t vs[3] = {vox_split ? vox_split[0] : 1024u, vox_split ? vox_split[1] : 512u,
                            vox_split ? vox_split[2] : 512u};
    const uint32_t ms[3] = {mat_split ? mat_split[0] : 512u, mat_split ? mat_split[1] : 512u,
                            mat_split ? mat_split[2] : 512u};
    if (!vs[1] || !vs[2] || !ms[0] || !ms[1] || !ms[2]) return VOX_ERR_INVALID;
    ctx->times = vox_stage_times{};
    Plan P;
    VOX_TRY(add_mesh(ctx, P, ctx->V, ctx->I, ctx->nt, ctx->mats, *g, kind, use_materials != 0, ms,
                     nodes_dev));
    VOX_TRY(run_plan(ctx, P, vs[1], vs[2]));
    return finish_times(ctx);
}

// ---- partitions and slices ----------------------------------------------------

namespace {
// rotateVertices (src/host_side.cpp:3261-3290): (x, y, z) ->
// (float(a - y), float(b + x), z), evaluated in double.
__global__ void k_rotate_vertices(const float *__restrict__ V, uint32_t nv, double a, double b,
                                  float *__restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
        const float x = V[3u * i], y = V[3u * i + 1u], z = V[3u * i + 2u];
        out[3u * i] = float(a - double(y));
        out[3u * i + 1u] = float(b + double(x));
        out[3u * i + 2u] = z;
    }
}

// unRotateNodes (src/voxelizer.cu:2061-2124) for Cartesian nodes of `size`
// bytes: node (x, y, z) of an X x Y x Z array goes to (y, X - 1 - x, z) of a
// Y x X x Z array.
__global__ void k_unrotate_nodes(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                 uint32_t X, uint32_t Y, uint32_t Z, uint32_t size) {
    const size_t n = (size_t)X * Y * Z;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % X), y = (uint32_t)((i / X) % Y);
        const size_t z = i / ((size_t)X * Y);
        const size_t o = (size_t)X * Y * z + (size_t)Y * (X - 1u - x) + y;
        for (uint32_t k = 0; k < size; k++) out[o * size + k] = in[i * size + k];
    }
}

// Bounding box of n device vertices (calculateBoundingBox).
vox_status device_bbox(vox_ctx *ctx, const float *V, uint32_t nv, double mn[3], double mx[3]) {
    uint32_t *bb;
    VOX_TRY(buf(ctx, "p_bbox", 6, &bb));
    VOX_TRY(fill_many(ctx, {{bb, 3, UINT32_MAX}, {bb + 3, 3, 0u}}));
    k_bbox<<<grid_for(nv, 256, 256), 256, 0, ctx->stream>>>(V, nv, bb);
    VOX_HIP(hipGetLastError());
    uint32_t h[6];
    VOX_TRY(read_u32(ctx, bb, 6, h));
    for (int c = 0; c < 3; c++) {
        uint32_t a = h[c], b = h[3 + c];
        uint32_t ua = (a & 0x80000000u) ? (a & 0x7FFFFFFFu) : ~a;
        uint32_t ub = (b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b;
        float fa, fb;
        std::memcpy(&fa, &ua, 4);
        std::memcpy(&fb, &ub, 4);
        mn[c] = fa;
        mx[c] = fb;
    }
    return VOX_OK;
}
}  // namespace

vox_status vox_voxelize_partition(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                                  int use_materials, const uint32_t vox_split[3],
                                  const uint32_t mat_split[3], const uint32_t part_min[2],
                                  const uint32_t part_max[2], void *nodes_dev) {
    if (!ctx || !grid_ok(g) || !nodes_dev || !part_min || !part_max) return VOX_ERR_INVALID;
    if (kind != VOX_NODE_SHORT && kind != VOX_NODE_LONG && kind != VOX_NODE_PARTIAL)
        return fail(ctx, VOX_ERR_UNSUPPORTED, "partitions are built for Cartesian nodes");
    for (int a = 0; a < 2; a++)
        if (part_min[a] >= part_max[a] || part_max[a] > g->res[1 + a])
            return fail(ctx, VOX_ERR_INVALID, "partition outside the grid");
    if (use_materials && !ctx->mats)
        return fail(ctx, VOX_ERR_INVALID, "materials requested but none set");
    const uint32_t vs[3] = {vox_split ? vox_split[0] : 1024u, vox_split ? vox_split[1] : 512u,
                            vox_split ? vox_split[2] : 512u};
    const uint32_t ms[3] = {mat_split ? mat_split[0] : 512u, mat_split ? mat_split[1] : 512u,
                            mat_split ? mat_split[2] : 512u};
    if (!vs[1] || !vs[2] || !ms[0] || !ms[1] || !ms[2]) return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    ctx->times = vox_stage_times{};
    const double zero[3] = {0.0, 0.0, 0.0};
    Plan P;
    VOX_TRY(add_job(ctx, P, ctx->V, ctx->I, ctx->nt, ctx->mats, *g, part_layout(*g, part_min, part_max),
                    zero, use_materials != 0, ms, ctx->cart_bdim, true, nodes_dev,
                    (int)kind, nullptr));
    VOX_TRY(run_plan(ctx, P, vs[1], vs[2]));
    return finish_times(ctx);
}

vox_status vox_solid_bits_partition(vox_ctx *ctx, const vox_grid *g, const uint32_t split_yz[2],
                                    const uint32_t part_min[2], const uint32_t part_max[2],
                                    uint32_t *bits_dev) {
    if (!ctx || !grid_ok(g) || !bits_dev || !part_min || !part_max) return VOX_ERR_INVALID;
    if (!ctx->V && ctx->nv) return fail(ctx, VOX_ERR_INVALID, "no mesh");
    for (int a = 0; a < 2; a++)
        if (part_min[a] >= part_max[a] || part_max[a] > g->res[1 + a])
            return fail(ctx, VOX_ERR_INVALID, "partition outside the grid");
    const uint32_t sy = split_yz ? split_yz[0] : 512u, sz = split_yz ? split_yz[1] : 512u;
    if (sy == 0 || sz == 0) return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    ctx->times = vox_stage_times{};
    const double zero[3] = {0.0, 0.0, 0.0};
    const uint32_t ms[3] = {512u, 512u, 512u};
    Plan P;
    VOX_TRY(add_job(ctx, P, ctx->V, ctx->I, ctx->nt, nullptr, *g,
                    part_layout(*g, part_min, part_max), zero, false, ms, ctx->cart_bdim, false,
                    nullptr, 0, bits_dev));
    PlanDev D;
    record(ctx, 0);
    VOX_TRY(plan_upload(ctx, P, D, sy, sz));
    VOX_TRY(records_stage(ctx, P, D));
    VOX_TRY(solid_stage(ctx, P, D));
    record(ctx, 1);
    if (ctx->timing) {
        VOX_HIP(hipEventSynchronize(ctx->ev[1]));
        (void)hipEventElapsedTime(&ctx->times.solid_fill_ms, ctx->ev[0], ctx->ev[1]);
        ctx->times.total_ms = ctx->times.solid_fill_ms;
    }
    return VOX_OK;
}

vox_status vox_voxelize_slice(vox_ctx *ctx, vox_plan_mode mode, double param, vox_node_kind kind,
                              int use_materials, int direction, uint32_t slice,
                              const uint32_t vox_split[2], const uint32_t mat_split[2],
                              void *nodes_dev, uint32_t dims_out[3]) {
    if (!ctx || !dims_out || direction < 0 || direction > 2) return VOX_ERR_INVALID;
    if (kind != VOX_NODE_SHORT && kind != VOX_NODE_LONG && kind != VOX_NODE_PARTIAL)
        return fail(ctx, VOX_ERR_UNSUPPORTED, "slices are built for Cartesian nodes");
    if (ctx->nv == 0) return fail(ctx, VOX_ERR_INVALID, "no vertices");
    // voxelizeSlice takes voxSplitRes and matSplitRes as (x, y); y serves z
    // too (src/host_side.cpp:1963-1970).
    const uint32_t vx = vox_split ? vox_split[0] : 1024u, vy = vox_split ? vox_split[1] : 512u;
    const uint32_t mx_ = mat_split ? mat_split[0] : 512u, my = mat_split ? mat_split[1] : 512u;
    const uint32_t ms[3] = {mx_, my, my};
    if (!vx || !vy || !mx_ || !my || !plan_args_ok(mode, param, vx)) return VOX_ERR_INVALID;
    if (use_materials && !ctx->mats)
        return fail(ctx, VOX_ERR_INVALID, "materials requested but none set");
    VOX_HIP(hipSetDevice(ctx->dev));
    ctx->times = vox_stage_times{};
    // The model: for slices along x, rotated a quarter turn about z first
    // (voxelizeEntry, src/host_side.cpp:703-716), around the centre
    // (max - min) / 2 of its bounding box.
    const float *V = ctx->V;
    double mn[3], mx[3];
    VOX_TRY(device_bbox(ctx, V, ctx->nv, mn, mx));
    if (direction == 0) {
        const double cx = (mx[0] - mn[0]) / 2.0, cy = (mx[1] - mn[1]) / 2.0;
        float *rv;
        VOX_TRY(buf(ctx, "slice_rot_v", (size_t)3 * ctx->nv, &rv));
        k_rotate_vertices<<<grid_for(ctx->nv, 256, 4096), 256, 0, ctx->stream>>>(
            V, ctx->nv, cy + cx, cy - cx, rv);
        VOX_HIP(hipGetLastError());
        V = rv;
        VOX_TRY(device_bbox(ctx, V, ctx->nv, mn, mx));
    }
    vox_grid g;
    plan_from_bbox(mn, mx, mode, param, vx, &g);
    if (!grid_ok(&g)) return fail(ctx, VOX_ERR_INVALID, "grid too large");
    // The slice (voxelizeEntry, 743-780): a one-row partition along y
    // (directions 0 and 1) or z; slice 0 and slices past the grid give zeros.
    const int a = direction == 2 ? 1 : 0;
    const bool oob = slice == 0 || slice > g.res[1 + a];
    const uint32_t srow = oob ? 0u : slice - 1u;
    uint32_t pmin[2] = {0u, 0u}, pmax[2] = {g.res[1], g.res[2]};
    pmin[a] = srow;
    pmax[a] = srow + 1u;
    Layout L = part_layout(g, pmin, pmax);
    const uint32_t X = g.alloc[0];
    if (direction == 0) {
        dims_out[0] = L.alloc[1];
        dims_out[1] = X;
    } else {
        dims_out[0] = X;
        dims_out[1] = L.alloc[1];
    }
    dims_out[2] = L.alloc[2];
    if (!nodes_dev) return VOX_OK;
    const size_t size = vox_node_size(kind);
    const size_t bytes = (size_t)dims_out[0] * dims_out[1] * dims_out[2] * size;
    if (oob) return fill(ctx, (uint32_t *)nodes_dev, bytes / 4, 0u);  // bytes: a multiple of 32
    const double zero[3] = {0.0, 0.0, 0.0};
    Plan P;
    if (direction == 0) {
        // The erosion runs on the un-rotated array in the reference, with
        // its y (the rotated x) split by voxSplitRes.y; computed here in the
        // rotated frame, which needs that axis whole.
        if (vy < X)
            return fail(ctx, VOX_ERR_UNSUPPORTED,
                        "x slices need voxSplitRes.y at least the rotated grid's x size");
        L.rot_ids = true;
        uint8_t *rot;
        VOX_TRY(buf(ctx, "slice_rot_nodes", bytes, &rot));
        VOX_TRY(add_job(ctx, P, V, ctx->I, ctx->nt, ctx->mats, g, L, zero, use_materials != 0, ms,
                        ctx->cart_bdim, true, rot, (int)kind, nullptr));
        // one erosion subspace across the rotated rows: vy >= X > the 3 rows
        VOX_TRY(run_plan(ctx, P, vy, vy));
        k_unrotate_nodes<<<grid_for(bytes / size, 256, 4096), 256, 0, ctx->stream>>>(
            rot, (uint8_t *)nodes_dev, X, L.alloc[1], L.alloc[2], (uint32_t)size);
        VOX_HIP(hipGetLastError());
    } else {
        VOX_TRY(add_job(ctx, P, V, ctx->I, ctx->nt, ctx->mats, g, L, zero, use_materials != 0, ms,
                        ctx->cart_bdim, true, nodes_dev, (int)kind, nullptr));
        VOX_TRY(run_plan(ctx, P, vy, vy));
    }
    return finish_times(ctx);
}

vox_status vox_enable_timing(vox_ctx *ctx, int enable) {
    if (!ctx) return VOX_ERR_INVALID;
    ctx->timing = enable != 0;
    return VOX_OK;
}

vox_status vox_get_stage_times(const vox_ctx *ctx, vox_stage_times *out) {
    if (!ctx || !out) return VOX_ERR_INVALID;
    *out = ctx->times;
    return VOX_OK;
}

// ---- surface brick shape ----------------------------------------------------

vox_status vox_set_brick_dims(vox_ctx *ctx, int fcc, const uint32_t dims[3]) {
    if (!ctx || !dims || !brick_ok(dims, fcc == 0))
        return ctx ? fail(ctx, VOX_ERR_INVALID, "unsupported brick shape") : VOX_ERR_INVALID;
    int *d = fcc ? ctx->fcc_bdim : ctx->cart_bdim;
    for (int a = 0; a < 3; a++) d[a] = (int)dims[a];
    return VOX_OK;
}

vox_status vox_get_brick_dims(const vox_ctx *ctx, int fcc, uint32_t dims[3]) {
    if (!ctx || !dims) return VOX_ERR_INVALID;
    const int *d = fcc ? ctx->fcc_bdim : ctx->cart_bdim;
    for (int a = 0; a < 3; a++) dims[a] = (uint32_t)d[a];
    return VOX_OK;
}

vox_status vox_autotune_bricks(vox_ctx *ctx, const vox_grid *grid, vox_node_kind kind,
                               const uint32_t vox_split[3], const uint32_t mat_split[3],
                               void *nodes_dev, uint32_t reps, uint32_t best_dims[3],
                               float *best_ms) {
    if (!ctx || !grid || !nodes_dev || reps == 0) return VOX_ERR_INVALID;
    const bool fcc = kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
    static const uint32_t cart[][3] = {{128, 4, 8}, {64, 4, 8}, {128, 2, 8}, {128, 4, 4},
                                       {256, 4, 4}, {256, 2, 8}, {64, 4, 4}, {128, 8, 4}};
    static const uint32_t fccs[][3] = {{64, 4, 4}, {64, 4, 8}, {128, 4, 4}, {64, 8, 4},
                                       {128, 2, 4}, {64, 2, 8}};
    const uint32_t(*cand)[3] = fcc ? fccs : cart;
    const size_t ncand = fcc ? sizeof(fccs) / sizeof(fccs[0]) : sizeof(cart) / sizeof(cart[0]);
    double best = 1e300;
    size_t bi = 0;
    for (size_t c = 0; c < ncand; c++) {
        VOX_TRY(vox_set_brick_dims(ctx, fcc ? 1 : 0, cand[c]));
        double t_best = 1e300;
        for (uint32_t r = 0; r <= reps; r++) {  // run 0 warms up
            VOX_TRY(vox_sync(ctx));
            const auto t0 = std::chrono::steady_clock::now();
            VOX_TRY(vox_voxelize_nodes(ctx, grid, kind, 1, vox_split, mat_split, nodes_dev));
            VOX_TRY(vox_sync(ctx));
            const double dt =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (r > 0) t_best = std::min(t_best, dt);
        }
        if (t_best < best) {
            best = t_best;
            bi = c;
        }
    }
    // Every shape writes the same nodes, so nodes_dev already holds them.
    VOX_TRY(vox_set_brick_dims(ctx, fcc ? 1 : 0, cand[bi]));
    if (best_dims)
        for (int a = 0; a < 3; a++) best_dims[a] = cand[bi][a];
    if (best_ms) *best_ms = (float)(best * 1e3);
    return VOX_OK;
}

// ---- batches ---------------------------------------------------------------

namespace {
vox_status run_item(vox_ctx *c, vox_batch_item &it) {
    VOX_TRY(vox_set_mesh_device(c, it.vertices_dev, it.nv, it.indices_dev, it.nt));
    VOX_TRY(vox_set_materials_device(c, it.materials_dev, it.n_unique));
    VOX_TRY(vox_voxelize_nodes(c, &it.grid, it.kind, it.materials_dev ? 1 : 0, it.vox_split,
                               it.mat_split, it.nodes_dev));
    return vox_sync(c);
}

bool item_ok(const vox_batch_item &it) {
    return grid_ok(&it.grid) && node_kind_ok(it.kind) && it.nodes_dev &&
           (it.nt == 0 || (it.vertices_dev && it.indices_dev)) && it.vox_split[1] &&
           it.vox_split[2] && it.mat_split[0] && it.mat_split[1] && it.mat_split[2];
}

// Meshes whose grid erodes as one y/z subspace (FCC lattices do not erode)
// can share the launches of a fused batch.
bool fusable(const vox_batch_item &it) {
    const bool fcc = it.kind == VOX_NODE_SHORT_FCC || it.kind == VOX_NODE_LONG_FCC;
    return fcc || (it.grid.alloc[1] <= it.vox_split[1] && it.grid.alloc[2] <= it.vox_split[2]);
}

// All listed meshes through one pipeline: every stage one launch.
vox_status fused_batch(vox_ctx *ctx, vox_batch_item *items, const std::vector<uint32_t> &idx) {
    if (idx.empty()) return VOX_OK;
    Plan P;
    for (uint32_t i : idx) {
        const vox_batch_item &it = items[i];
        VOX_TRY(add_mesh(ctx, P, it.vertices_dev, it.indices_dev, it.nt, it.materials_dev,
                         it.grid, it.kind, it.materials_dev != nullptr, it.mat_split,
                         it.nodes_dev));
    }
    ctx->times = vox_stage_times{};
    VOX_TRY(run_plan(ctx, P, UINT32_MAX, UINT32_MAX));
    VOX_TRY(finish_times(ctx));
    return vox_sync(ctx);
}
}  // namespace

vox_status vox_voxelize_batch(vox_ctx *ctx, vox_batch_item *items, uint32_t n,
                              vox_schedule schedule, uint32_t streams) {
    if (!ctx || (n && !items)) return VOX_ERR_INVALID;
    if (schedule != VOX_SCHEDULE_PLAIN && schedule != VOX_SCHEDULE_LEARNED) return VOX_ERR_INVALID;
    for (uint32_t i = 0; i < n; i++) {
        items[i].status = item_ok(items[i]) ? VOX_OK : VOX_ERR_INVALID;
        if (items[i].status != VOX_OK) {
            ctx->err = "batch item " + std::to_string(i) + " is invalid";
            return VOX_ERR_INVALID;
        }
    }
    VOX_HIP(hipSetDevice(ctx->dev));
    vox_status first = VOX_OK;
    const int saved = ctx->schedule;
    if (streams <= 1) {
        ctx->schedule = schedule;
        std::vector<uint32_t> fused, single;
        for (uint32_t i = 0; i < n; i++)
            (streams == 0 && fusable(items[i]) ? fused : single).push_back(i);
        const vox_status st = fused_batch(ctx, items, fused);
        for (uint32_t i : fused) items[i].status = st;
        if (st != VOX_OK) first = st;
        for (uint32_t i : single) {
            items[i].status = run_item(ctx, items[i]);
            if (items[i].status != VOX_OK && first == VOX_OK) first = items[i].status;
        }
        ctx->schedule = saved;
        if (first != VOX_OK && st == VOX_OK) ctx->err = "a batch item failed";
        return first;
    }
    streams = std::min<uint32_t>(streams, 64);
    while (ctx->workers.size() < streams) {
        vox_ctx *w = nullptr;
        const vox_status st = vox_create(ctx->dev, &w);
        if (st != VOX_OK) return fail(ctx, st, "cannot create a worker context");
        ctx->workers.push_back(w);
    }
    std::vector<float> params;
    if (ctx->cost_ready) {
        params.resize(kCostParams);
        VOX_TRY(vox_cost_get_params(ctx, params.data()));
    }
    for (uint32_t k = 0; k < streams; k++) {
        vox_ctx *w = ctx->workers[k];
        w->schedule = schedule;
        w->heavy_cost = ctx->heavy_cost;
        w->screen_ratio = ctx->screen_ratio;
        w->split_parts = ctx->split_parts;
        w->timing = false;
        for (int a = 0; a < 3; a++) {
            w->cart_bdim[a] = ctx->cart_bdim[a];
            w->fcc_bdim[a] = ctx->fcc_bdim[a];
        }
        if (!params.empty()) VOX_TRY(vox_cost_set_params(w, params.data()));
    }
    // Longest estimated work first.
    std::vector<uint32_t> order(n);
    std::vector<double> est(n);
    for (uint32_t i = 0; i < n; i++) {
        order[i] = i;
        const size_t nodes = vox_node_count(&items[i].grid, items[i].kind);
        est[i] = (double)nodes * (double)vox_node_size(items[i].kind) + 64.0 * items[i].nt;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return est[a] > est[b]; });
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> threads;
    for (uint32_t k = 0; k < streams; k++)
        threads.emplace_back([&, k]() {
            vox_ctx *w = ctx->workers[k];
            for (uint32_t q = next.fetch_add(1); q < n; q = next.fetch_add(1)) {
                vox_batch_item &it = items[order[q]];
                it.status = run_item(w, it);
            }
        });
    for (auto &t : threads) t.join();
    for (uint32_t i = 0; i < n; i++)
        if (items[i].status != VOX_OK) {
            first = items[i].status;
            ctx->err = "batch item " + std::to_string(i) + " failed";
            break;
        }
    return first;
}

// ---- replayed batches -------------------------------------------------------

struct vox_graph {
    vox_ctx *w = nullptr;  // private context: stream, buffers, job table
    std::vector<vox_batch_item> items;
    int schedule = VOX_SCHEDULE_PLAIN;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const uint32_t *changed = nullptr;  // last erosion pass's change flag (or null)
    uint32_t fallbacks = 0;
};

namespace {
void drop_graph(vox_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
}

vox_status graph_plan(vox_graph *g, Plan &P) {
    for (const vox_batch_item &it : g->items)
        VOX_TRY(add_mesh(g->w, P, it.vertices_dev, it.indices_dev, it.nt, it.materials_dev, it.grid,
                         it.kind, it.materials_dev != nullptr, it.mat_split, it.nodes_dev));
    return VOX_OK;
}

// Erosion subspaces: a single mesh keeps its voxSplitRes y/z; a batch's
// meshes each erode whole (they are fusable).
uint32_t graph_split(const vox_graph *g, int a) {
    return g->items.size() == 1 ? g->items[0].vox_split[a] : UINT32_MAX;
}

// One run without the graph (it also records the sizes the run needed).
vox_status graph_eager(vox_graph *g) {
    vox_ctx *w = g->w;
    w->fixed = false;
    w->schedule = g->schedule;
    Plan P;
    VOX_TRY(graph_plan(g, P));
    VOX_TRY(run_plan(w, P, graph_split(g, 1), graph_split(g, 2)));
    return vox_sync(w);
}

// Captures the stages with the sizes of the last eager run.
vox_status graph_capture(vox_graph *g) {
    vox_ctx *w = g->w;
    drop_graph(g);
    Plan P;
    VOX_TRY(graph_plan(g, P));
    const uint32_t nj = (uint32_t)P.jobs.size();
    w->cap_solid = w->last_solid + w->last_solid / 8 + 256;
    w->cap_surf = w->last_surf + w->last_surf / 8 + 256;
    w->n_passes = w->last_passes + 1;
    // Every buffer the capture uses is sized now: nothing may allocate or
    // synchronise while the stream is captured.
    uint32_t *u;
    uint2 *pr;
    VOX_TRY(buf(w, "g_ovf", 1, &w->ovf));
    VOX_TRY(buf(w, "s_pairs", w->cap_solid, &pr));
    VOX_TRY(buf(w, "s_list", w->cap_solid, &u));
    uint16_t *u16;
    VOX_TRY(buf(w, "s_pmask", w->cap_solid, &u16));
    VOX_TRY(buf(w, "s_list_mask", w->cap_solid, &u16));
    VOX_TRY(buf(w, "f_blist", w->cap_surf, &u));
    VOX_TRY(buf(w, "s_first", block_starts_size(w->cap_solid), &u));
    VOX_TRY(buf(w, "f_first", block_starts_size(w->cap_surf), &u));
    VOX_TRY(buf(w, "e_changes", (size_t)w->n_passes * nj, &u));
    VOX_TRY(buf(w, "e_totals", w->n_passes, &u));
    VOX_TRY(buf(w, "e_notconv", 1, &u));
    PlanDev D;
    VOX_TRY(plan_upload(w, P, D, graph_split(g, 1), graph_split(g, 2)));
    VOX_TRY(vox_sync(w));
    w->fixed = true;
    const bool timing = w->timing;
    w->timing = false;
    VOX_HIP_G(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    const vox_status st = run_stages(w, P, D, graph_split(g, 1), graph_split(g, 2));
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(w->stream, &graph);
    w->fixed = false;
    w->timing = timing;
    if (st != VOX_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    VOX_HIP_G(ec);
    g->graph = graph;
    g->changed = w->last_changed;
    VOX_HIP_G(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return VOX_OK;
}
}  // namespace

vox_status vox_graph_create(vox_ctx *ctx, const vox_batch_item *items, uint32_t n,
                            vox_schedule schedule, vox_graph **out) {
    if (!ctx || !out || n == 0 || !items) return VOX_ERR_INVALID;
    if (schedule != VOX_SCHEDULE_PLAIN && schedule != VOX_SCHEDULE_LEARNED) return VOX_ERR_INVALID;
    *out = nullptr;
    for (uint32_t i = 0; i < n; i++) {
        if (!item_ok(items[i])) return fail(ctx, VOX_ERR_INVALID, "graph item " + std::to_string(i) + " is invalid");
        if (n > 1 && !fusable(items[i]))
            return fail(ctx, VOX_ERR_UNSUPPORTED,
                        "graph item " + std::to_string(i) + " erodes in several subspaces");
    }
    vox_graph *g = new vox_graph();
    g->items.assign(items, items + n);
    g->schedule = schedule;
    vox_status st = vox_create(ctx->dev, &g->w);
    if (st != VOX_OK) {
        delete g;
        return fail(ctx, st, "cannot create the graph's context");
    }
    vox_ctx *w = g->w;
    w->heavy_cost = ctx->heavy_cost;
    w->screen_ratio = ctx->screen_ratio;
    w->split_parts = ctx->split_parts;
    w->embed_limit = ctx->embed_limit;
    for (int a = 0; a < 3; a++) {
        w->cart_bdim[a] = ctx->cart_bdim[a];
        w->fcc_bdim[a] = ctx->fcc_bdim[a];
    }
    if (ctx->cost_ready) {
        std::vector<float> params(kCostParams);
        st = vox_cost_get_params(ctx, params.data());
        if (st == VOX_OK) st = vox_cost_set_params(w, params.data());
    }
    if (st == VOX_OK) st = graph_eager(g);
    if (st == VOX_OK) st = graph_capture(g);
    if (st != VOX_OK) {
        ctx->err = "graph: " + w->err;
        vox_graph_destroy(g);
        return st;
    }
    *out = g;
    return VOX_OK;
}

vox_status vox_graph_launch(vox_graph *g) {
    if (!g || !g->exec) return VOX_ERR_INVALID;
    vox_ctx *w = g->w;
    VOX_HIP_G(hipSetDevice(w->dev));
    VOX_HIP_G(hipGraphLaunch(g->exec, w->stream));
    uint32_t ovf = 0, changed = 0;
    if (read_u32(w, w->ovf, 1, &ovf) != VOX_OK) return VOX_ERR_HIP;
    if (g->changed && read_u32(w, g->changed, 1, &changed) != VOX_OK) return VOX_ERR_HIP;
    if (!ovf && !changed) return VOX_OK;
    // The lists were too short or erosion had not converged: the meshes need
    // more than the graph was sized for.  Run without it and capture anew.
    g->fallbacks++;
    vox_status st = graph_eager(g);
    if (st == VOX_OK) st = graph_capture(g);
    return st;
}

uint32_t vox_graph_fallbacks(const vox_graph *g) { return g ? g->fallbacks : 0u; }

void vox_graph_destroy(vox_graph *g) {
    if (!g) return;
    drop_graph(g);
    if (g->w) vox_destroy(g->w);
    delete g;
}

// ---- brick schedule and cost model ----------------------------------------

vox_status vox_set_schedule(vox_ctx *ctx, vox_schedule schedule) {
    if (!ctx || (schedule != VOX_SCHEDULE_PLAIN && schedule != VOX_SCHEDULE_LEARNED))
        return VOX_ERR_INVALID;
    ctx->schedule = schedule;
    return VOX_OK;
}

vox_status vox_set_heavy_threshold(vox_ctx *ctx, float threshold) {
    if (!ctx || threshold != threshold) return VOX_ERR_INVALID;
    ctx->heavy_cost = threshold;
    return VOX_OK;
}

vox_status vox_set_screen_ratio(vox_ctx *ctx, float ratio) {
    if (!ctx || !(ratio >= 0.0f)) return VOX_ERR_INVALID;
    ctx->screen_ratio = ratio;
    return VOX_OK;
}

vox_status vox_set_split_parts(vox_ctx *ctx, uint32_t parts) {
    if (!ctx || (parts != 0u && parts != 1u && parts != 2u && parts != 4u))
        return VOX_ERR_INVALID;
    ctx->split_parts = parts;
    return VOX_OK;
}

vox_status vox_cost_set_params(vox_ctx *ctx, const float *params) {
    if (!ctx || !params) return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    if (!ctx->cost_params) VOX_HIP(hipMalloc(&ctx->cost_params, kCostParams * sizeof(float)));
    VOX_HIP(hipMemcpyAsync(ctx->cost_params, params, kCostParams * sizeof(float),
                           hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->cost_ready = true;
    return VOX_OK;
}

vox_status vox_cost_get_params(vox_ctx *ctx, float *params) {
    if (!ctx || !params) return VOX_ERR_INVALID;
    if (!ctx->cost_ready) return fail(ctx, VOX_ERR_INVALID, "no cost model set");
    VOX_HIP(hipMemcpyAsync(params, ctx->cost_params, kCostParams * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    return VOX_OK;
}

vox_status vox_cost_record(vox_ctx *ctx, int enable) {
    if (!ctx) return VOX_ERR_INVALID;
    ctx->record = enable != 0;
    if (ctx->record) {
        ctx->rec_feats.clear();
        ctx->rec_cycles.clear();
    }
    return VOX_OK;
}

vox_status vox_cost_samples(vox_ctx *ctx, float *features, float *cycles, uint32_t cap,
                            uint32_t *n_out) {
    if (!ctx) return VOX_ERR_INVALID;
    const uint32_t n = (uint32_t)ctx->rec_cycles.size();
    if (n_out) *n_out = n;
    const uint32_t m = std::min(n, cap);
    if (features)
        std::memcpy(features, ctx->rec_feats.data(), (size_t)m * kCostFeatures * sizeof(float));
    if (cycles) std::memcpy(cycles, ctx->rec_cycles.data(), (size_t)m * sizeof(float));
    return VOX_OK;
}

vox_status vox_mlp_forward(vox_ctx *ctx, const float *params_dev, const float *x_dev, uint32_t n,
                           float *y_dev) {
    if (!ctx || !params_dev || !x_dev || !y_dev) return VOX_ERR_INVALID;
    if (n == 0) return VOX_OK;
    VOX_HIP(hipSetDevice(ctx->dev));
    k_mlp_forward<<<(n + kMlpBlockRows - 1) / kMlpBlockRows, 256, 0, ctx->stream>>>(
        params_dev, x_dev, n, y_dev);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

namespace {
// Gradient of (1/n) sum (f(x) - t)^2 over n rows into grad (device); the sum
// of squared errors into loss_dev (device, may be null).
vox_status mlp_grad(vox_ctx *ctx, const float *P, const float *X, const float *T, uint32_t n,
                    float *grad, float *loss_dev) {
    const uint32_t nblocks = (n + kMlpBlockRows - 1) / kMlpBlockRows;
    float *partial, *ploss;
    VOX_TRY(buf(ctx, "m_partial", (size_t)nblocks * kCostParams, &partial));
    VOX_TRY(buf(ctx, "m_ploss", nblocks, &ploss));
    k_mlp_grad<<<nblocks, 256, 0, ctx->stream>>>(P, X, T, n, 1.0f / (float)n, partial, ploss);
    VOX_HIP(hipGetLastError());
    k_mlp_reduce<<<(kCostParams + 1 + 255) / 256, 256, 0, ctx->stream>>>(partial, ploss, nblocks,
                                                                           grad, loss_dev);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// splitmix64: deterministic stream for initialisation and shuffling.
uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
}  // namespace

vox_status vox_mlp_grad(vox_ctx *ctx, const float *params_dev, const float *x_dev,
                        const float *t_dev, uint32_t n, float *grad_dev, float *loss) {
    if (!ctx || !params_dev || !x_dev || !t_dev || !grad_dev || n == 0) return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    float *ldev;
    VOX_TRY(buf(ctx, "m_loss", 1, &ldev));
    VOX_TRY(mlp_grad(ctx, params_dev, x_dev, t_dev, n, grad_dev, ldev));
    uint32_t bits_;
    VOX_TRY(read_u32(ctx, (const uint32_t *)ldev, 1, &bits_));
    float l;
    std::memcpy(&l, &bits_, 4);
    if (loss) *loss = l / (float)n;
    return VOX_OK;
}

vox_status vox_cost_train(vox_ctx *ctx, const float *features, const float *cycles, uint32_t n,
                          uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                          float *final_mse) {
    if (!ctx || !features || !cycles || n < kMlpBlockRows || epochs == 0 || batch == 0 ||
        batch % kMlpBlockRows != 0 || !(lr > 0.0f))
        return VOX_ERR_INVALID;
    VOX_HIP(hipSetDevice(ctx->dev));
    // Standardized log-cycle targets; rows shuffled once (deterministically).
    double mean = 0.0, var = 0.0;
    std::vector<double> y(n);
    for (uint32_t i = 0; i < n; i++) {
        y[i] = std::log(std::max<double>(cycles[i], 1.0));
        mean += y[i];
    }
    mean /= n;
    for (uint32_t i = 0; i < n; i++) var += (y[i] - mean) * (y[i] - mean);
    const double sd = std::max(std::sqrt(var / n), 1e-6);
    uint64_t rs = 0x5EED0000ull + seed;
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    for (uint32_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[splitmix64(rs) % (i + 1)]);
    std::vector<float> X((size_t)n * kCostFeatures), T(n);
    for (uint32_t i = 0; i < n; i++) {
        std::memcpy(&X[(size_t)i * kCostFeatures], features + (size_t)perm[i] * kCostFeatures,
                    kCostFeatures * sizeof(float));
        T[i] = (float)((y[perm[i]] - mean) / sd);
    }
    // He-uniform weights, zero biases.
    std::vector<float> P(kCostParams, 0.0f);
    auto init = [&](int off, int count, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int k = 0; k < count; k++) P[off + k] = (float)((2.0 * uniform01(rs) - 1.0) * a);
    };
    init(0, 64 * 16, 16);
    init(1088, 64 * 64, 64);
    init(5248, 64, 64);
    float *dX, *dT, *dP, *dG, *dM, *dV, *dL;
    VOX_TRY(buf(ctx, "m_x", X.size(), &dX));
    VOX_TRY(buf(ctx, "m_t", n, &dT));
    VOX_TRY(buf(ctx, "m_p", kCostParams, &dP));
    VOX_TRY(buf(ctx, "m_g", kCostParams, &dG));
    VOX_TRY(buf(ctx, "m_m", kCostParams, &dM));
    VOX_TRY(buf(ctx, "m_v", kCostParams, &dV));
    VOX_TRY(buf(ctx, "m_loss", 1, &dL));
    VOX_HIP(hipMemcpyAsync(dX, X.data(), X.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipMemcpyAsync(dT, T.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    VOX_HIP(hipMemcpyAsync(dP, P.data(), kCostParams * 4, hipMemcpyHostToDevice, ctx->stream));
    VOX_TRY(fill(ctx, (uint32_t *)dM, kCostParams, 0u));
    VOX_TRY(fill(ctx, (uint32_t *)dV, kCostParams, 0u));
    const float b1 = 0.9f, b2 = 0.999f;
    double p1 = 1.0, p2 = 1.0;
    for (uint32_t e = 0; e < epochs; e++)
        for (uint32_t r0 = 0; r0 < n; r0 += batch) {
            const uint32_t rows = std::min(batch, n - r0);
            VOX_TRY(mlp_grad(ctx, dP, dX + (size_t)r0 * kCostFeatures, dT + r0, rows, dG,
                             nullptr));
            p1 *= b1;
            p2 *= b2;
            k_adam<<<(kCostParams + 255) / 256, 256, 0, ctx->stream>>>(
                dP, dG, dM, dV, kCostParams, lr, b1, b2, 1e-8f, (float)(1.0 - p1),
                (float)(1.0 - p2));
            VOX_HIP(hipGetLastError());
        }
    VOX_TRY(mlp_grad(ctx, dP, dX, dT, n, dG, dL));
    uint32_t lb;
    VOX_TRY(read_u32(ctx, (const uint32_t *)dL, 1, &lb));
    float l;
    std::memcpy(&l, &lb, 4);
    if (final_mse) *final_mse = l / (float)n;
    if (!ctx->cost_params) VOX_HIP(hipMalloc(&ctx->cost_params, kCostParams * sizeof(float)));
    VOX_HIP(hipMemcpyAsync(ctx->cost_params, dP, kCostParams * 4, hipMemcpyDeviceToDevice,
                           ctx->stream));
    VOX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->cost_ready = true;
    return VOX_OK;
}

}  // extern "C"
