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
#include <cstring>
#include <map>
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
    uint32_t mat_slot_cap = 1u << 16;  // bytes per material arena (multiple of 4)
    // persistent surface launch: compute units and resident bricks per CU
    int n_cu = 0;
    int surf_blocks_per_cu = 0;
    // brick schedule of the surface pass and its cost model
    int schedule = VOX_SCHEDULE_PLAIN;
    float *cost_params = nullptr;  // device, kCostParams floats
    bool cost_ready = false;
    float heavy_cost = 1.5f;    // standardized log cost above which a brick goes first
    float screen_ratio = 1.5f;  // pair count / mean above which the model is consulted
    bool record = false;
    std::vector<float> rec_feats, rec_cycles;
    // worker contexts of batched runs (same device, own streams and buffers)
    std::vector<vox_ctx *> workers;
    // per-triangle vertex records of the current call (k_tri_records)
    float4 *rec = nullptr;
    // surface brick shapes (voxels along x, y, z)
    int cart_bdim[3] = {kCartBX, kCartBY, kCartBZ};
    int fcc_bdim[3] = {kFccBX, kFccBY, kFccBZ};
};

#define VOX_HIP(expr)                                                               \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ctx->err = std::string(hipGetErrorString(e_)) + " @ " + #expr;          \
            return e_ == hipErrorOutOfMemory ? VOX_ERR_OOM : VOX_ERR_HIP;           \
        }                                                                           \
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

template <class T>
vox_status buf(vox_ctx *ctx, const char *name, size_t count, T **out) {
    Buf &b = ctx->bufs[name];
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (bytes > b.cap) {
        if (b.p) {
            VOX_HIP(hipStreamSynchronize(ctx->stream));
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

// Row-major launches: block (64, 4) covers 4 rows of up to 64 words; a 1D
// grid padded to a multiple of the XCD count (see xcd_logical_block).
inline dim3 row_grid(uint32_t W, uint32_t Y, uint32_t Z) {
    return dim3(xcd_grid(((W + kWave - 1) / kWave) * ((Y + 3) / 4) * Z));
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

// Exclusive scan of in[0,n) into out[0,n]; out[n] is the total.
vox_status scan(vox_ctx *ctx, const uint32_t *in, uint32_t *out, uint32_t n, int depth = 0) {
    const uint32_t nblocks = n / kScanTile + 1;
    char name[32];
    std::snprintf(name, sizeof(name), "scan_part%d", depth);
    uint32_t *part;
    VOX_TRY(buf(ctx, name, nblocks + 1, &part));
    k_scan_blocks<<<nblocks, kScanThreads, 0, ctx->stream>>>(in, out, n, part);
    VOX_HIP(hipGetLastError());
    if (nblocks > 1) {
        std::snprintf(name, sizeof(name), "scan_psum%d", depth);
        uint32_t *psum;
        VOX_TRY(buf(ctx, name, nblocks + 1, &psum));
        VOX_TRY(scan(ctx, part, psum, nblocks, depth + 1));
        k_scan_add<<<nblocks, kScanThreads, 0, ctx->stream>>>(out, n, psum);
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
// multiple of 64, at most 32 rows (y*z), at most kMaxBrickVox voxels, and (for
// the Cartesian node pass, whose threads own kNodeZ planes) z a multiple of
// kNodeZ.
bool brick_ok(const uint32_t b[3], bool cart) {
    for (int a = 0; a < 3; a++)
        if (b[a] == 0 || (b[a] & (b[a] - 1)) != 0) return false;
    return b[0] >= 64 && b[1] * b[2] <= 32 && b[0] * b[1] * b[2] <= (uint32_t)kMaxBrickVox &&
           (!cart || b[2] % kNodeZ == 0);
}

// Per-triangle vertex records of the current mesh, read by the solid and
// surface stages instead of the indexed vertex arrays.
vox_status make_records(vox_ctx *ctx) {
    VOX_TRY(buf(ctx, "tri_rec", (size_t)std::max<uint32_t>(ctx->nt, 1) * 3, &ctx->rec));
    if (ctx->nt) {
        k_tri_records<<<grid_for(ctx->nt, 256), 256, 0, ctx->stream>>>(ctx->V, ctx->I, ctx->nt,
                                                                        ctx->rec);
        VOX_HIP(hipGetLastError());
    }
    return VOX_OK;
}

void record(vox_ctx *ctx, int i) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

// ---------------------------------------------------------------------------
// Solid stage for one grid (possibly FCC-shifted) over its y/z subspaces.
// ---------------------------------------------------------------------------
vox_status solid_grid(vox_ctx *ctx, const vox_grid &g, const double extMin[3], uint32_t splitY,
                      uint32_t splitZ, uint32_t *bits, const char *tag) {
    const uint32_t X = g.alloc[0], Yext = g.res[1], Zext = g.res[2];
    const uint32_t W = X / 32;
    if (W == 0 || (size_t)16 * W * 4 > 65536)
        return fail(ctx, VOX_ERR_UNSUPPORTED, "x resolution beyond the fill kernel's LDS row buffer");
    if (g.x_splits == 0 || X / g.x_splits > 1024)
        return fail(ctx, VOX_ERR_UNSUPPORTED,
                    "voxSplitRes.x gives x splits wider than 1024 voxels, where the reference's "
                    "word bitmask is undefined");
    const uint32_t rowsY = g.alloc[1], rowsZ = g.alloc[2];
    const uint32_t tilesY = (Yext + 3) / 4, tilesZ = (Zext + 3) / 4;
    const uint32_t nTiles = tilesY * tilesZ;
    const uint32_t nt = ctx->nt;
    (void)tag;

    uint32_t *cand_count, *cand_off, *tile_hist, *tile_off, *tile_cursor, *list, *counters;
    uint2 *pairs;
    VOX_TRY(buf(ctx, "s_cand_count", nt, &cand_count));
    VOX_TRY(buf(ctx, "s_cand_off", nt + 1, &cand_off));
    VOX_TRY(buf(ctx, "s_tile_hist", nTiles, &tile_hist));
    VOX_TRY(buf(ctx, "s_tile_off", nTiles + 1, &tile_off));
    VOX_TRY(buf(ctx, "s_tile_cursor", nTiles, &tile_cursor));
    VOX_TRY(buf(ctx, "s_counters", 4, &counters));
    uint4 *cbox;
    VOX_TRY(buf(ctx, "s_cbox", nt, &cbox));

    std::vector<Span> ys = split_axis_max(Yext, splitY), zs = split_axis_max(Zext, splitZ);
    for (const Span &zsp : zs)
        for (const Span &ysp : ys) {
            SolidParams p;
            p.V = ctx->V;
            p.I = ctx->I;
            p.rec = ctx->rec;
            p.nt = nt;
            p.emx = extMin[0];
            p.emy = extMin[1];
            p.emz = extMin[2];
            p.d = g.voxel_length;
            p.smin[0] = ysp.min;
            p.smin[1] = zsp.min;
            p.smax[0] = ysp.max;
            p.smax[1] = zsp.max;
            p.tilesY = tilesY;
            p.X = X;
            p.rowsY = rowsY;
            if (nt > 0) {
                k_solid_setup<<<grid_for(nt, 256), 256, 0, ctx->stream>>>(p, cand_count, cbox);
                VOX_HIP(hipGetLastError());
            }
            VOX_TRY(scan(ctx, cand_count, cand_off, nt));
            uint32_t total = 0;
            VOX_TRY(read_u32(ctx, cand_off + nt, 1, &total));
            VOX_TRY(buf(ctx, "s_pairs", total, &pairs));
            VOX_TRY(buf(ctx, "s_list", total, &list));
            VOX_TRY(fill(ctx, counters, 4, 0u));
            VOX_TRY(fill(ctx, tile_hist, nTiles, 0u));
            VOX_TRY(fill(ctx, tile_cursor, nTiles, 0u));
            if (total > 0) {
                k_solid_candidates<<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                    p, cand_off, cand_off + nt, cbox, pairs, tile_hist);
                VOX_HIP(hipGetLastError());
            }
            VOX_TRY(scan(ctx, tile_hist, tile_off, nTiles));
            if (total > 0) {
                k_solid_scatter<<<grid_for(total, 256), 256, 0, ctx->stream>>>(
                    pairs, cand_off + nt, tile_off, tile_cursor, list);
                VOX_HIP(hipGetLastError());
            }
            const uint32_t ty0 = ysp.min / 4, ty1 = (ysp.max - 1) / 4;
            const uint32_t tz0 = zsp.min / 4, tz1 = (zsp.max - 1) / 4;
            const uint32_t nty = ty1 - ty0 + 1, ntz = tz1 - tz0 + 1;
            if (ysp.max > ysp.min && zsp.max > zsp.min) {
                k_solid_fill<<<nty * ntz, kWave, 16 * W * 4, ctx->stream>>>(
                    p, tile_off, tile_hist, list, ty0, nty, tz0, bits);
                VOX_HIP(hipGetLastError());
            }
            ctx->times.solid_pairs += total;  // candidate (triangle, tile) pairs
        }
    k_zero_pad_rows<<<grid_for((size_t)(2 * rowsZ + 2 * rowsY) * W, 256, 4096), 256, 0,
                      ctx->stream>>>(bits, W, rowsY, rowsZ);
    VOX_HIP(hipGetLastError());
    return VOX_OK;
}

// ---------------------------------------------------------------------------
// Surface stage for one grid.
// ---------------------------------------------------------------------------
struct SurfOut {
    uint32_t *surf;
    uint32_t *brick_mat_off;
    uint16_t *wrank;
    uint8_t *mat_list;
    uint32_t nb[3];
};

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

vox_status surface_grid(vox_ctx *ctx, const vox_grid &g, const double extMin[3],
                        const uint32_t matSplit[3], const int bdim[3], uint32_t *bits,
                        const char *tag, SurfOut &out) {
    const uint32_t nt = ctx->nt;
    SurfParams p;
    p.V = ctx->V;
    p.I = ctx->I;
    p.nt = nt;
    p.m = F3{(float)extMin[0], (float)extMin[1], (float)extMin[2]};
    p.d = (float)g.voxel_length;
    for (int a = 0; a < 3; a++) {
        p.res[a] = (int32_t)g.res[a];
        p.alloc[a] = g.alloc[a];
        p.bdim[a] = bdim[a];
        p.nb[a] = (g.alloc[a] + bdim[a] - 1) / bdim[a];
        out.nb[a] = p.nb[a];
    }
    const uint32_t nbt = p.nb[0] * p.nb[1] * p.nb[2];
    std::string t(tag);
    uint4 *tri_info, *tri_ext = nullptr;
    uint32_t *cand, *cand_off, *brick_hist, *brick_off, *brick_cursor, *blist, *stats;
    SplitTable *dsplit;
    const bool learned = ctx->schedule == VOX_SCHEDULE_LEARNED && ctx->cost_ready;
    const bool features = learned || ctx->record;
    if (ctx->timing) ctx->times.learned_schedule = learned ? 1u : 0u;
    VOX_TRY(buf(ctx, (t + "_tri_info").c_str(), nt, &tri_info));
    if (features) VOX_TRY(buf(ctx, "f_tri_ext", nt, &tri_ext));
    float4 *tri_rec = ctx->rec;
    VOX_TRY(buf(ctx, "f_cand", nt, &cand));
    VOX_TRY(buf(ctx, "f_cand_off", nt + 1, &cand_off));
    VOX_TRY(buf(ctx, "f_brick_hist", nbt, &brick_hist));
    VOX_TRY(buf(ctx, "f_brick_off", nbt + 1, &brick_off));
    VOX_TRY(buf(ctx, "f_brick_cursor", nbt, &brick_cursor));
    VOX_TRY(buf(ctx, "f_stats", kMatSlots * kSlotStride + 1, &stats));
    VOX_TRY(buf(ctx, (t + "_split").c_str(), 1, &dsplit));
    VOX_TRY(buf(ctx, (t + "_surf").c_str(), (size_t)g.alloc[0] / 32 * g.alloc[1] * g.alloc[2],
                &out.surf));
    VOX_TRY(buf(ctx, (t + "_bmo").c_str(), nbt, &out.brick_mat_off));
    const uint32_t nwords = (uint32_t)(bdim[0] * bdim[1] * bdim[2] / 32);
    VOX_TRY(buf(ctx, (t + "_wrank").c_str(), (size_t)nbt * nwords, &out.wrank));

    // matSplitRes subspaces over extPartition (src/host_side.cpp:1107-1111).
    SplitTable st;
    std::memset(&st, 0, sizeof(st));
    for (int a = 0; a < 3; a++) {
        std::vector<Span> sp = split_axis_max(g.res[a], matSplit[a]);
        if (sp.empty() || sp.size() > (size_t)kMaxSplits)
            return fail(ctx, VOX_ERR_UNSUPPORTED, "matSplitRes gives too many subspaces");
        st.n[a] = (uint32_t)sp.size();
        for (size_t k = 0; k < sp.size(); k++) {
            st.lo[a][k] = sp[k].min;
            st.hi[a][k] = sp[k].max;
        }
    }
    VOX_HIP(hipMemcpyAsync(dsplit, &st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));

    if (nt > 0) {
        k_surf_setup<<<grid_for(nt, 256), 256, 0, ctx->stream>>>(p, tri_info, cand, tri_ext,
                                                                 tri_rec);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(scan(ctx, cand, cand_off, nt));
    uint32_t total = 0;
    VOX_TRY(read_u32(ctx, cand_off + nt, 1, &total));
    VOX_TRY(buf(ctx, "f_blist", total, &blist));
    VOX_TRY(fill(ctx, brick_hist, nbt, 0u));
    VOX_TRY(fill(ctx, brick_cursor, nbt, 0u));
    if (total > 0) {
        k_surf_bin<<<grid_for(total, 256), 256, 0, ctx->stream>>>(
            p, cand_off, cand_off + nt, tri_info, brick_hist, brick_off, brick_cursor, blist, 0);
        VOX_HIP(hipGetLastError());
    }
    VOX_TRY(scan(ctx, brick_hist, brick_off, nbt));
    // Non-empty bricks in index order (the plain schedule of the brick pass).
    uint32_t *flags, *blist_order, *blist_order_off, *next;
    VOX_TRY(buf(ctx, "f_flags", nbt, &flags));
    VOX_TRY(buf(ctx, "f_order", nbt, &blist_order));
    VOX_TRY(buf(ctx, "f_order_off", nbt + 1, &blist_order_off));
    VOX_TRY(buf(ctx, "f_next", 1, &next));
    k_nonzero_flags<<<grid_for(nbt, 256), 256, 0, ctx->stream>>>(brick_hist, flags, nbt);
    VOX_HIP(hipGetLastError());
    VOX_TRY(scan(ctx, flags, blist_order_off, nbt));
    k_nonzero_list<<<grid_for(nbt, 256), 256, 0, ctx->stream>>>(brick_hist, blist_order_off,
                                                                blist_order, nbt);
    VOX_HIP(hipGetLastError());
    if (total > 0) {
        k_surf_bin<<<grid_for(total, 256), 256, 0, ctx->stream>>>(
            p, cand_off, cand_off + nt, tri_info, brick_hist, brick_off, brick_cursor, blist, 1);
        VOX_HIP(hipGetLastError());
    }
    if (ctx->timing) ctx->times.surface_pairs += total;
    // Learned schedule: the listed bricks sorted by predicted cost, most
    // expensive first (features -> MLP -> 1024-bucket counting sort).
    const uint32_t *n_list = blist_order_off + nbt;
    const uint32_t *brick_list = blist_order;
    float *feats = nullptr;
    uint64_t *cycles = nullptr;
    if (features) VOX_TRY(buf(ctx, "f_feats", (size_t)nbt * kCostFeatures, &feats));
    if (ctx->record) {  // training samples: features of every listed brick
        k_brick_features<<<grid_for(nbt, 4, 65535), 256, 0, ctx->stream>>>(
            blist_order, nullptr, n_list, brick_hist, brick_off, blist, tri_info, tri_ext, bdim[0],
            bdim[1], bdim[2], feats);
        VOX_HIP(hipGetLastError());
    } else if (learned) {
        // Screen by pair count, predict the candidates' cost, order.
        float *pred;
        uint32_t *flags2, *soff, *sel, *light, *loff, *chist, *coff, *ccur, *sorted;
        VOX_TRY(buf(ctx, "f_pred", nbt, &pred));
        VOX_TRY(buf(ctx, "f_sflags", nbt, &flags2));
        VOX_TRY(buf(ctx, "f_soff", nbt + 1, &soff));
        VOX_TRY(buf(ctx, "f_sel", nbt, &sel));
        VOX_TRY(buf(ctx, "f_light", nbt, &light));
        VOX_TRY(buf(ctx, "f_loff", nbt + 1, &loff));
        VOX_TRY(buf(ctx, "f_chist", kCostBuckets, &chist));
        VOX_TRY(buf(ctx, "f_coff", kCostBuckets + 1, &coff));
        VOX_TRY(buf(ctx, "f_ccur", kCostBuckets, &ccur));
        VOX_TRY(buf(ctx, "f_sorted", nbt, &sorted));
        VOX_TRY(fill(ctx, flags2, nbt, 0u));
        VOX_TRY(fill(ctx, light, nbt, 0u));
        VOX_TRY(fill(ctx, chist, kCostBuckets, 0u));
        VOX_TRY(fill(ctx, ccur, kCostBuckets, 0u));
        const uint32_t g256 = grid_for(nbt, 256, 1024);
        k_cost_screen<<<g256, 256, 0, ctx->stream>>>(blist_order, n_list, brick_hist,
                                                     cand_off + nt, ctx->screen_ratio, flags2,
                                                     light);
        VOX_HIP(hipGetLastError());
        VOX_TRY(scan(ctx, flags2, soff, nbt));
        const uint32_t *n_sel = soff + nbt;
        k_cost_compact<<<g256, 256, 0, ctx->stream>>>(flags2, soff, n_list, sel);
        VOX_HIP(hipGetLastError());
        k_brick_features<<<grid_for(nbt, 4, 65535), 256, 0, ctx->stream>>>(
            blist_order, sel, n_sel, brick_hist, brick_off, blist, tri_info, tri_ext, bdim[0],
            bdim[1], bdim[2], feats);
        VOX_HIP(hipGetLastError());
        k_mlp_forward_n<<<(nbt + kMlpBlockRows - 1) / kMlpBlockRows, 256, 0, ctx->stream>>>(
            ctx->cost_params, feats, n_sel, pred);
        VOX_HIP(hipGetLastError());
        k_cost_hist<<<g256, 256, 0, ctx->stream>>>(pred, sel, n_sel, ctx->heavy_cost, light,
                                                   chist);
        VOX_HIP(hipGetLastError());
        VOX_TRY(scan(ctx, chist, coff, kCostBuckets));
        VOX_TRY(scan(ctx, light, loff, nbt));
        k_cost_scatter<<<g256, 256, 0, ctx->stream>>>(pred, blist_order, n_list, sel, n_sel,
                                                      ctx->heavy_cost, coff, light, loff, ccur,
                                                      sorted);
        VOX_HIP(hipGetLastError());
        brick_list = sorted;
    }
    if (ctx->record) VOX_TRY(buf(ctx, "f_cycles", nbt, &cycles));
    for (int attempt = 0; attempt < 3; attempt++) {
        const uint32_t slot_cap = ctx->mat_slot_cap;
        VOX_TRY(buf(ctx, (t + "_mats").c_str(), (size_t)slot_cap * kMatSlots, &out.mat_list));
        VOX_TRY(fill(ctx, stats, kMatSlots * kSlotStride + 1, 0u));
        uint32_t *overflow = stats + kMatSlots * kSlotStride;
        // Empty bricks are not visited: mark them up front.
        VOX_TRY(fill(ctx, out.brick_mat_off, nbt, UINT32_MAX));
        VOX_TRY(fill(ctx, next, 1, 0u));
        const uint32_t blocks =
            std::min<uint32_t>(nbt, (uint32_t)(ctx->n_cu * ctx->surf_blocks_per_cu));
        k_surf_brick<<<blocks, 256, 0, ctx->stream>>>(
            p, dsplit, tri_rec, brick_off, brick_hist, blist, ctx->mats, bits, out.surf,
            out.brick_mat_off, out.wrank, out.mat_list, stats, slot_cap, overflow, brick_list,
            n_list, next, cycles);
        VOX_HIP(hipGetLastError());
        std::vector<uint32_t> h(kMatSlots * kSlotStride + 1);
        VOX_TRY(read_u32(ctx, stats, h.size(), h.data()));
        uint32_t used = 0;
        uint64_t vox = 0;
        for (uint32_t k = 0; k < kMatSlots; k++) {
            used = std::max(used, h[k * kSlotStride]);
            vox += h[k * kSlotStride + 1];
        }
        if (h.back() == 0u) {
            if (ctx->timing) ctx->times.surface_voxels += (uint32_t)vox;
            if (ctx->record) VOX_TRY(record_samples(ctx, blist_order, n_list, feats, cycles, nbt));
            return VOX_OK;
        }
        ctx->mat_slot_cap = ((used + used / 4 + 1024u) + 3u) & ~3u;
    }
    return fail(ctx, VOX_ERR_OOM, "surface material list overflow");
}

// Erosion to the fixed point, per alloc y/z subspace in the reference's order,
// in place on `bits`.  Passes are launched kBatch at a time; a pass after one
// that changed nothing exits at once, so one host read per batch suffices.
vox_status erode(vox_ctx *ctx, const GridDims &gd, uint32_t splitY, uint32_t splitZ,
                 uint32_t *bits) {
    constexpr uint32_t kBatch = 4;
    const uint32_t nzc = (gd.Z + kErodeZ - 1) / kErodeZ;
    const uint32_t nblocks = ((gd.W + kWave - 1) / kWave) * ((gd.Y + 3) / 4) * nzc;
    uint32_t *changes, *stamp;
    VOX_TRY(buf(ctx, "e_changes", kBatch, &changes));
    VOX_TRY(buf(ctx, "e_stamp", nblocks, &stamp));
    VOX_TRY(fill(ctx, stamp, nblocks, 0u));
    std::vector<Span> ys = split_axis_max(gd.Y, splitY), zs = split_axis_max(gd.Z, splitZ);
    uint32_t passes = 0, gpass = 0;
    for (const Span &zsp : zs)
        for (const Span &ysp : ys) {
            bool first = true;
            for (int guard = 0; guard < 1000000; guard++) {
                VOX_TRY(fill(ctx, changes, kBatch, 0u));
                for (uint32_t k = 0; k < kBatch; k++) {
                    k_erode<<<row_grid(gd.W, gd.Y, nzc), dim3(kWave, 4), 0, ctx->stream>>>(
                        gd, bits, ysp.min, ysp.max, zsp.min, zsp.max, stamp, changes, k,
                        gpass + k, (first && k == 0) ? 1 : 0);
                    VOX_HIP(hipGetLastError());
                }
                gpass += kBatch;
                first = false;
                uint32_t h[kBatch];
                VOX_TRY(read_u32(ctx, changes, kBatch, h));
                uint32_t k0 = kBatch;
                for (uint32_t k = 0; k < kBatch; k++)
                    if (h[k] == 0u) {
                        k0 = k;
                        break;
                    }
                if (k0 < kBatch) {
                    passes += k0 + 1;
                    break;
                }
                passes += kBatch;
            }
        }
    ctx->times.erosion_passes += passes;
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
        hipHostMalloc((void **)&ctx->pinned, kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return VOX_ERR_HIP;
    }
    for (auto &e : ctx->ev) (void)hipEventCreate(&e);
    if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&ctx->surf_blocks_per_cu, k_surf_brick, 256,
                                                     0) != hipSuccess) {
        vox_destroy(ctx);
        return VOX_ERR_HIP;
    }
    ctx->n_cu = std::max(ctx->n_cu, 1);
    ctx->surf_blocks_per_cu = std::max(ctx->surf_blocks_per_cu, 1);
    uint8_t lut[64] = {0};
    for (auto &pr : kOrientPairs) lut[pr[0]] = pr[1];
    if (set_orient_table(lut) != hipSuccess) {
        vox_destroy(ctx);
        return VOX_ERR_HIP;
    }
    *out = ctx;
    return VOX_OK;
}

void vox_destroy(vox_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->dev);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
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
    if (ctx->cost_params) (void)hipFree(ctx->cost_params);
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
    VOX_TRY(fill(ctx, bb, 3, UINT32_MAX));
    VOX_TRY(fill(ctx, bb + 3, 3, 0u));
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
    double em[3];
    for (int a = 0; a < 3; a++) em[a] = g->min_vertex[a] + (shift ? shift[a] : 0.0);
    const uint32_t sy = split_yz ? split_yz[0] : 512u, sz = split_yz ? split_yz[1] : 512u;
    if (sy == 0 || sz == 0) return VOX_ERR_INVALID;
    record(ctx, 0);
    VOX_TRY(make_records(ctx));
    VOX_TRY(solid_grid(ctx, *g, em, sy, sz, bits_dev, "g1"));
    record(ctx, 1);
    if (ctx->timing) {
        VOX_HIP(hipEventSynchronize(ctx->ev[1]));
        (void)hipEventElapsedTime(&ctx->times.solid_fill_ms, ctx->ev[0], ctx->ev[1]);
        ctx->times.total_ms = ctx->times.solid_fill_ms;
    }
    return VOX_OK;
}

vox_status vox_voxelize_nodes(vox_ctx *ctx, const vox_grid *g, vox_node_kind kind,
                              int use_materials, const uint32_t vox_split[3],
                              const uint32_t mat_split[3], void *nodes_dev) {
    if (!ctx || !grid_ok(g) || !nodes_dev) return VOX_ERR_INVALID;
    if (kind == VOX_NODE_PARTIAL)
        return fail(ctx, VOX_ERR_UNSUPPORTED, "PartialNode (volume ratio) is not built");
    const bool fcc = kind == VOX_NODE_SHORT_FCC || kind == VOX_NODE_LONG_FCC;
    if (kind != VOX_NODE_SHORT && kind != VOX_NODE_LONG && !fcc) return VOX_ERR_INVALID;
    if (use_materials && !ctx->mats)
        return fail(ctx, VOX_ERR_INVALID, "materials requested but none set");
    VOX_HIP(hipSetDevice(ctx->dev));
    const uint32_t vs[3] = {vox_split ? vox_split[0] : 1024u, vox_split ? vox_split[1] : 512u,
                            vox_split ? vox_split[2] : 512u};
    const uint32_t ms[3] = {mat_split ? mat_split[0] : 512u, mat_split ? mat_split[1] : 512u,
                            mat_split ? mat_split[2] : 512u};
    if (!vs[1] || !vs[2] || !ms[0] || !ms[1] || !ms[2]) return VOX_ERR_INVALID;
    ctx->times = vox_stage_times{};
    const size_t words = (size_t)g->alloc[0] / 32 * g->alloc[1] * g->alloc[2];
    const double base[3] = {g->min_vertex[0], g->min_vertex[1], g->min_vertex[2]};
    GridDims gd{g->alloc[0], g->alloc[1], g->alloc[2], g->alloc[0] / 32};
    record(ctx, 0);
    VOX_TRY(make_records(ctx));
    if (!fcc) {
        uint32_t *bits;
        VOX_TRY(buf(ctx, "n_bits", words, &bits));
        VOX_TRY(solid_grid(ctx, *g, base, vs[1], vs[2], bits, "g1"));
        record(ctx, 1);
        SurfOut so{};
        const int *bdim = ctx->cart_bdim;
        if (use_materials) VOX_TRY(surface_grid(ctx, *g, base, ms, bdim, bits, "g1", so));
        record(ctx, 2);
        // Several erosion subspaces: keep the pre-erosion grid for the ids
        // of nodes next to a later subspace (see k_nodes_cart).
        std::vector<Span> eys = split_axis_max(g->alloc[1], vs[1]);
        std::vector<Span> ezs = split_axis_max(g->alloc[2], vs[2]);
        uint32_t *pre = nullptr;
        uint8_t *ystart = nullptr, *zstart = nullptr;
        if (eys.size() > 1 || ezs.size() > 1) {
            VOX_TRY(buf(ctx, "n_bits_pre", words, &pre));
            VOX_HIP(hipMemcpyAsync(pre, bits, words * 4, hipMemcpyDeviceToDevice, ctx->stream));
            std::vector<uint8_t> fy(g->alloc[1], 0), fz(g->alloc[2], 0);
            for (size_t k = 1; k < eys.size(); k++) fy[eys[k].min] = 1;
            for (size_t k = 1; k < ezs.size(); k++) fz[ezs[k].min] = 1;
            VOX_TRY(buf(ctx, "n_ystart", fy.size(), &ystart));
            VOX_TRY(buf(ctx, "n_zstart", fz.size(), &zstart));
            VOX_HIP(hipMemcpyAsync(ystart, fy.data(), fy.size(), hipMemcpyHostToDevice, ctx->stream));
            VOX_HIP(hipMemcpyAsync(zstart, fz.data(), fz.size(), hipMemcpyHostToDevice, ctx->stream));
            VOX_HIP(hipStreamSynchronize(ctx->stream));
        }
        VOX_TRY(erode(ctx, gd, vs[1], vs[2], bits));
        record(ctx, 3);
        MatSource msrc{so.surf, so.brick_mat_off, so.wrank, so.mat_list};
        k_nodes_cart<<<row_grid(gd.W, gd.Y, (gd.Z + kNodeZ - 1) / kNodeZ), dim3(kWave, 4), 0,
                       ctx->stream>>>(
            gd, bits, pre, ystart, zstart, msrc, use_materials ? 1 : 0,
            kind == VOX_NODE_SHORT ? 0 : 1, nodes_dev, brick_dims(*g, bdim));
        VOX_HIP(hipGetLastError());
        record(ctx, 4);
    } else {
        // fccWorker (src/host_side.cpp:1262-1698): four sub-grids shifted by
        // (0,0,0), (h,h,0), (0,h,h), (h,0,h), h = d/2 (1436-1463).
        const double h = g->voxel_length / 2.0;
        const double shifts[4][3] = {{0.0, 0.0, 0.0}, {h, h, 0.0}, {0.0, h, h}, {h, 0.0, h}};
        uint32_t *gb[4];
        SurfOut so[4] = {};
        const char *tags[4] = {"fg1", "fg2", "fg3", "fg4"};
        char name[32];
        for (int k = 0; k < 4; k++) {
            std::snprintf(name, sizeof(name), "n_fcc_bits%d", k);
            VOX_TRY(buf(ctx, name, words, &gb[k]));
            const double em[3] = {base[0] + shifts[k][0], base[1] + shifts[k][1],
                                  base[2] + shifts[k][2]};
            VOX_TRY(solid_grid(ctx, *g, em, vs[1], vs[2], gb[k], tags[k]));
        }
        record(ctx, 1);
        const int *bdim = ctx->fcc_bdim;
        if (use_materials) {
            for (int k = 0; k < 4; k++) {
                const double em[3] = {base[0] + shifts[k][0], base[1] + shifts[k][1],
                                      base[2] + shifts[k][2]};
                VOX_TRY(surface_grid(ctx, *g, em, ms, bdim, gb[k], tags[k], so[k]));
            }
        }
        record(ctx, 2);
        record(ctx, 3);
        MatSource m[4];
        for (int k = 0; k < 4; k++)
            m[k] = MatSource{so[k].surf, so[k].brick_mat_off, so[k].wrank, so[k].mat_list};
        k_nodes_fcc<<<row_grid(2 * gd.W, gd.Y, 2 * gd.Z), dim3(kWave, 4), 0, ctx->stream>>>(
            gd, gb[0], gb[1], gb[2], gb[3], m[0], m[1], m[2], m[3], use_materials ? 1 : 0,
            kind == VOX_NODE_SHORT_FCC ? 3 : 4, nodes_dev, brick_dims(*g, bdim));
        VOX_HIP(hipGetLastError());
        record(ctx, 4);
    }
    if (ctx->timing) {
        VOX_HIP(hipEventSynchronize(ctx->ev[4]));
        (void)hipEventElapsedTime(&ctx->times.solid_fill_ms, ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&ctx->times.surface_ms, ctx->ev[1], ctx->ev[2]);
        (void)hipEventElapsedTime(&ctx->times.erode_ms, ctx->ev[2], ctx->ev[3]);
        (void)hipEventElapsedTime(&ctx->times.nodes_ms, ctx->ev[3], ctx->ev[4]);
        (void)hipEventElapsedTime(&ctx->times.total_ms, ctx->ev[0], ctx->ev[4]);
    }
    return VOX_OK;
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
}  // namespace

vox_status vox_voxelize_batch(vox_ctx *ctx, vox_batch_item *items, uint32_t n,
                              vox_schedule schedule, uint32_t streams) {
    if (!ctx || (n && !items)) return VOX_ERR_INVALID;
    if (schedule != VOX_SCHEDULE_PLAIN && schedule != VOX_SCHEDULE_LEARNED) return VOX_ERR_INVALID;
    vox_status first = VOX_OK;
    if (schedule == VOX_SCHEDULE_PLAIN || streams <= 1) {
        const int saved = ctx->schedule;
        ctx->schedule = schedule;
        for (uint32_t i = 0; i < n; i++) {
            items[i].status = run_item(ctx, items[i]);
            if (items[i].status != VOX_OK && first == VOX_OK) first = items[i].status;
        }
        ctx->schedule = saved;
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
        w->schedule = VOX_SCHEDULE_LEARNED;
        w->heavy_cost = ctx->heavy_cost;
        w->timing = false;
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
