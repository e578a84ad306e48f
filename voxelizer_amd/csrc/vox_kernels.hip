// vox_kernels.hip -- HIP kernels of the voxelizer hot path for gfx950.
// The arithmetic of every test follows the reference expressions cited in
// vox_math.h; the data flow (binning, LDS staging, bitwise node stencils) is
// this library's own.
#include "vox_kernels.h"

namespace vk {

// ---------------------------------------------------------------------------
// Scan
// ---------------------------------------------------------------------------
// Exclusive scan of in[0, n) into out[0, n] (out[n] = total); one block scans
// kScanTile elements and leaves its sum in partials[block].
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const uint32_t *__restrict__ in,
                                                              uint32_t *__restrict__ out,
                                                              uint32_t n,
                                                              uint32_t *__restrict__ partials) {
    __shared__ uint32_t sh[kScanThreads];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)tid * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        uint64_t idx = base + k;
        v[k] = idx < n ? in[idx] : 0u;
        s += v[k];
    }
    sh[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        uint32_t t = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    uint32_t excl = tid ? sh[tid - 1] : 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        uint64_t idx = base + k;
        if (idx <= n) out[idx] = excl;
        excl += v[k];
    }
    if (tid == kScanThreads - 1) partials[blockIdx.x] = sh[tid];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t *__restrict__ out,
                                                           uint32_t n,
                                                           const uint32_t *__restrict__ partials) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    const uint32_t add = partials[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < (uint32_t)kScanTile; k += kScanThreads) {
        uint64_t idx = base + k;
        if (idx <= n) out[idx] += add;
    }
}

// ---------------------------------------------------------------------------
// Bounding box: order-preserving uint keys of the float coordinates.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_bbox(const float *__restrict__ V, uint32_t nv,
                                              uint32_t *__restrict__ out6) {
    uint32_t mn[3] = {UINT32_MAX, UINT32_MAX, UINT32_MAX}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            uint32_t k = fkey(V[3u * i + c]);
            mn[c] = min(mn[c], k);
            mx[c] = max(mx[c], k);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[c] = min(mn[c], (uint32_t)__shfl_xor((int)mn[c], off));
            mx[c] = max(mx[c], (uint32_t)__shfl_xor((int)mx[c], off));
        }
    }
    // Block reduction through LDS, then one set of atomics per block.
    __shared__ uint32_t red[6][256 / kWave];
    const uint32_t wv = threadIdx.x / kWave;
    if (lane_id() == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            red[c][wv] = mn[c];
            red[3 + c][wv] = mx[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        uint32_t r = red[c][0];
        for (uint32_t k = 1; k < blockDim.x / kWave; k++)
            r = c < 3 ? min(r, red[c][k]) : max(r, red[c][k]);
        if (c < 3)
            atomicMin(&out6[c], r);
        else
            atomicMax(&out6[c], r);
    }
}

// ---------------------------------------------------------------------------
// Solid stage
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_tri_d(const SolidParams &p, uint32_t t, D3 v[3]) {
    v[0] = fetch_d(p.V, p.I, t, 0);
    v[1] = fetch_d(p.V, p.I, t, 1);
    v[2] = fetch_d(p.V, p.I, t, 2);
}

// Candidate tiles of each triangle: its commonSpace tile rectangle.
__global__ __launch_bounds__(256) void k_solid_setup(SolidParams p,
                                                     uint32_t *__restrict__ cand_count) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < p.nt; t += gridDim.x * blockDim.x) {
        D3 v[3];
        load_tri_d(p, t, v);
        TileSetup ts;
        uint32_t c = 0;
        if (tile_setup(v, p.emy, p.emz, p.d, p.smin[0], p.smin[1], p.smax[0], p.smax[1], ts))
            c = (ts.tmax[0] - ts.tmin[0] + 1u) * (ts.tmax[1] - ts.tmin[1] + 1u);
        cand_count[t] = c;
    }
}

// One thread per candidate (triangle, tile): the reference's per-tile queue
// test, done in parallel instead of one thread walking all tiles of a
// triangle.  Queued pairs are appended with a wave-aggregated bump and
// counted per tile for the counting sort.
__global__ __launch_bounds__(256) void k_solid_candidates(
    SolidParams p, const uint32_t *__restrict__ cand_off, const uint32_t *__restrict__ cand_total,
    uint2 *__restrict__ pairs, uint32_t *__restrict__ pair_count, uint32_t pair_cap,
    uint32_t *__restrict__ tile_hist) {
    const uint32_t total = *cand_total;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        const uint32_t c = base + threadIdx.x;
        bool hit = false;
        uint32_t t = 0, tile = 0;
        if (c < total) {
            t = upper_bound_u32(cand_off, p.nt, c) - 1u;
            D3 v[3];
            load_tri_d(p, t, v);
            TileSetup ts;
            if (tile_setup(v, p.emy, p.emz, p.d, p.smin[0], p.smin[1], p.smax[0], p.smax[1], ts)) {
                const uint32_t w = ts.tmax[0] - ts.tmin[0] + 1u;
                const uint32_t local = c - cand_off[t];
                const uint32_t ty = ts.tmin[0] + local % w, tz = ts.tmin[1] + local / w;
                hit = tile_hit(ts, ty, tz, p.emy, p.emz, p.d);
                tile = tz * p.tilesY + ty;
            }
        }
        const uint32_t pos = wave_bump(pair_count, hit);
        if (hit) {
            if (pos < pair_cap) pairs[pos] = make_uint2(t, tile);
            atomicAdd(&tile_hist[tile], 1u);
        }
    }
}

__global__ __launch_bounds__(256) void k_solid_scatter(const uint2 *__restrict__ pairs,
                                                       const uint32_t *__restrict__ pair_count,
                                                       const uint32_t *__restrict__ tile_off,
                                                       uint32_t *__restrict__ tile_cursor,
                                                       uint32_t *__restrict__ list) {
    const uint32_t n = *pair_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint2 pr = pairs[i];
        uint32_t pos = tile_off[pr.y] + atomicAdd(&tile_cursor[pr.y], 1u);
        list[pos] = pr.x;
    }
}

// One wave per 4x4 tile of voxel-centre rows.  Every queued triangle's
// crossings toggle one bit each in an LDS copy of the tile's 16 rows
// (ds_xor); a per-row prefix XOR then turns crossings into solid runs
// (the suffix flips of src/voxelizer.cu:1487-1616) and each bit word is
// written to HBM once, coalesced, with no global atomics.
__global__ __launch_bounds__(kWave) void k_solid_fill(SolidParams p,
                                                      const uint32_t *__restrict__ tile_off,
                                                      const uint32_t *__restrict__ tile_hist,
                                                      const uint32_t *__restrict__ list,
                                                      uint32_t ty0, uint32_t nty, uint32_t tz0,
                                                      uint32_t *__restrict__ bits) {
    extern __shared__ uint32_t cross[];  // [16][W]
    const uint32_t W = p.X >> 5;
    const uint32_t tid = threadIdx.x;
    const uint32_t ty = ty0 + blockIdx.x % nty;
    const uint32_t tz = tz0 + blockIdx.x / nty;
    for (uint32_t i = tid; i < 16u * W; i += kWave) cross[i] = 0u;
    __syncthreads();
    const uint32_t tile = tz * p.tilesY + ty;
    const uint32_t cnt = tile_hist[tile];
    const uint32_t off = tile_off[tile];
    for (uint32_t i = tid; i < cnt; i += kWave) {
        const uint32_t t = list[off + i];
        D3 v[3];
        load_tri_d(p, t, v);
        FillSetup fs;
        fill_setup(v, fs);
        for (uint32_t r = 0; r < 16u; r++) {
            const uint32_t cy = ty * 4u + (r & 3u), cz = tz * 4u + (r >> 2);
            if (cy < p.smin[0] || cy >= p.smax[0] || cz < p.smin[1] || cz >= p.smax[1]) continue;
            uint32_t vx;
            if (fill_center(fs, p.emx, p.emy, p.emz, p.d, cy, cz, vx) && vx < p.X)
                atomicXor(&cross[r * W + (vx >> 5)], 1u << (31u - (vx & 31u)));
        }
    }
    __syncthreads();
    // Crossing parity -> solid bits (MSB-first bits: x%32 == 0 is bit 31).
    if (tid < 16u) {
        uint32_t carry = 0u;
        uint32_t *row = cross + tid * W;
        for (uint32_t w = 0; w < W; w++) {
            uint32_t c = row[w];
            uint32_t s = c;
            s ^= s >> 1;
            s ^= s >> 2;
            s ^= s >> 4;
            s ^= s >> 8;
            s ^= s >> 16;
            row[w] = s ^ carry;
            carry ^= (__popc(c) & 1u) ? 0xFFFFFFFFu : 0u;
        }
    }
    __syncthreads();
    for (uint32_t r = 0; r < 16u; r++) {
        const uint32_t cy = ty * 4u + (r & 3u), cz = tz * 4u + (r >> 2);
        if (cy < p.smin[0] || cy >= p.smax[0] || cz < p.smin[1] || cz >= p.smax[1]) continue;
        const size_t rowIdx = (size_t)(cz + 1u) * p.rowsY + (cy + 1u);
        for (uint32_t w = tid; w < W; w += kWave) bits[rowIdx * W + w] = cross[r * W + w];
    }
}

__global__ __launch_bounds__(256) void k_zero_pad_rows(uint32_t *__restrict__ bits,
                                                       uint32_t W, uint32_t rowsY,
                                                       uint32_t rowsZ) {
    const uint64_t nrows = 2ull * rowsZ + 2ull * rowsY;
    const uint64_t total = nrows * W;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t r = i / W, w = i % W;
        uint32_t y, z;
        if (r < 2ull * rowsZ) {
            z = (uint32_t)(r >> 1);
            y = (r & 1ull) ? rowsY - 1u : 0u;
        } else {
            uint64_t q = r - 2ull * rowsZ;
            y = (uint32_t)(q >> 1);
            z = (q & 1ull) ? rowsZ - 1u : 0u;
        }
        bits[((uint64_t)z * rowsY + y) * W + w] = 0u;
    }
}

// ---------------------------------------------------------------------------
// Surface stage
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_tri_f(const SurfParams &p, uint32_t t, F3 tri[3]) {
    tri[0] = fetch_f(p.V, p.I, t, 0);
    tri[1] = fetch_f(p.V, p.I, t, 1);
    tri[2] = fetch_f(p.V, p.I, t, 2);
}

// Voxel range of a triangle clipped to the whole ext space (the union of all
// matSplitRes subspaces); false if it lies outside (adjustVoxelRange,
// src/voxelizer.cu:3963-3993).
__device__ __forceinline__ bool surf_range(const SurfParams &p, const F3 tri[3], VoxBB &vb) {
    F3 mn, mx;
    tri_bounds(tri, mn, mx);
    vb = voxel_bounds_half(mn, mx, p.m, p.d);
    if (0 > vb.max.x || 0 > vb.max.y || 0 > vb.max.z) return false;
    if (vb.min.x > p.res[0] - 1 || vb.min.y > p.res[1] - 1 || vb.min.z > p.res[2] - 1) return false;
    vb.min = {max(vb.min.x, 0), max(vb.min.y, 0), max(vb.min.z, 0)};
    vb.max = {min(vb.max.x, p.res[0] - 1), min(vb.max.y, p.res[1] - 1), min(vb.max.z, p.res[2] - 1)};
    return true;
}

// Bricks are laid over alloc coordinates (ext + 1).
__device__ __forceinline__ void brick_range(const SurfParams &p, const VoxBB &vb, uint32_t lo[3],
                                            uint32_t hi[3]) {
    lo[0] = (uint32_t)(vb.min.x + 1) / (uint32_t)p.bdim[0];
    lo[1] = (uint32_t)(vb.min.y + 1) / (uint32_t)p.bdim[1];
    lo[2] = (uint32_t)(vb.min.z + 1) / (uint32_t)p.bdim[2];
    hi[0] = (uint32_t)(vb.max.x + 1) / (uint32_t)p.bdim[0];
    hi[1] = (uint32_t)(vb.max.y + 1) / (uint32_t)p.bdim[1];
    hi[2] = (uint32_t)(vb.max.z + 1) / (uint32_t)p.bdim[2];
}

// Per triangle: class and dominant axis (classifyTriangles), and the box of
// bricks its clipped voxel range touches, packed as
// {info, lo.x | lo.y<<16, lo.z | w.x<<16, w.y | w.z<<16}.
__global__ __launch_bounds__(256) void k_surf_setup(SurfParams p, uint4 *__restrict__ tri_info,
                                                    uint32_t *__restrict__ cand_count) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < p.nt; t += gridDim.x * blockDim.x) {
        F3 tri[3];
        load_tri_f(p, t, tri);
        int bb, axis;
        classify(tri, p.m, p.d, bb, axis);
        uint32_t c = 0;
        VoxBB vb;
        uint32_t lo[3] = {0, 0, 0}, w[3] = {0, 0, 0};
        if (bb != BB_DEGEN && surf_range(p, tri, vb)) {
            uint32_t hi[3];
            brick_range(p, vb, lo, hi);
            for (int a = 0; a < 3; a++) w[a] = hi[a] - lo[a] + 1u;
            c = w[0] * w[1] * w[2];
        }
        const uint32_t info = ((uint32_t)bb << 30) | ((uint32_t)axis << 28);
        tri_info[t] = make_uint4(info, lo[0] | (lo[1] << 16), lo[2] | (w[0] << 16),
                                 w[1] | (w[2] << 16));
        cand_count[t] = c;
    }
}

// One thread per (triangle, brick) candidate: counting pass (scatter = 0)
// or placement pass (scatter = 1) of the counting sort by brick.  The
// triangle of candidate c is found by binary search in the candidate offsets,
// narrowed per block to the triangles its 256 candidates span.
__global__ __launch_bounds__(256) void k_surf_bin(
    SurfParams p, const uint32_t *__restrict__ cand_off, const uint32_t *__restrict__ cand_total,
    const uint4 *__restrict__ tri_info, uint32_t *__restrict__ brick_hist,
    const uint32_t *__restrict__ brick_off, uint32_t *__restrict__ brick_cursor,
    uint32_t *__restrict__ blist, int scatter) {
    __shared__ uint32_t trange[2];
    const uint32_t total = *cand_total;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        __syncthreads();
        if (threadIdx.x < 2) {
            const uint32_t key = threadIdx.x == 0 ? base : min(base + blockDim.x, total) - 1u;
            trange[threadIdx.x] = upper_bound_u32(cand_off, p.nt, key) - 1u;
        }
        __syncthreads();
        const uint32_t c = base + threadIdx.x;
        if (c >= total) continue;
        uint32_t lo = trange[0], hi = trange[1] + 1u;
        while (lo + 1u < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cand_off[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint32_t t = lo;
        const uint4 ti = tri_info[t];
        const uint32_t bx0 = ti.y & 0xFFFFu, by0 = ti.y >> 16, bz0 = ti.z & 0xFFFFu;
        const uint32_t w0 = ti.z >> 16, w1 = ti.w & 0xFFFFu;
        const uint32_t local = c - cand_off[t];
        const uint32_t bx = bx0 + local % w0;
        const uint32_t by = by0 + (local / w0) % w1;
        const uint32_t bz = bz0 + local / (w0 * w1);
        const uint32_t b = (bz * p.nb[1] + by) * p.nb[0] + bx;
        if (!scatter) {
            atomicAdd(&brick_hist[b], 1u);
        } else {
            uint32_t pos = brick_off[b] + atomicAdd(&brick_cursor[b], 1u);
            blist[pos] = t;
        }
    }
}

__device__ __forceinline__ int32_t split_hi_of(const SplitTable &s, int axis, int32_t i,
                                               int32_t &lo_out) {
    const uint32_t n = s.n[axis];
    for (uint32_t k = 0; k < n; k++) {
        if ((uint32_t)i < s.hi[axis][k]) {
            lo_out = (int32_t)s.lo[axis][k];
            return (int32_t)s.hi[axis][k];
        }
    }
    lo_out = (int32_t)s.lo[axis][n - 1];
    return (int32_t)s.hi[axis][n - 1];
}

struct BrickBox {
    int32_t lo[3], hi[3];  // inclusive, ext coordinates
};

__device__ __forceinline__ void lds_write(uint32_t *keys, const BrickBox &bb, const int32_t bd[3],
                                          int32_t x, int32_t y, int32_t z, uint32_t key) {
    if (x < bb.lo[0] || x > bb.hi[0] || y < bb.lo[1] || y > bb.hi[1] || z < bb.lo[2] ||
        z > bb.hi[2])
        return;
    const int32_t lx = x - bb.lo[0], ly = y - bb.lo[1], lz = z - bb.lo[2];
    atomicMax(&keys[(lz * bd[1] + ly) * bd[0] + lx], key);
}

// Per-triangle data of one brick's chunk, set up once in LDS.
struct SurfTri {
    OverlapData o[3];          // projections YZ (0), ZX (1), XY (2)
    F3 n, t0;                  // normal, first vertex
    int32_t vmin[3], vmax[3];  // voxel range clipped to the ext space
    int32_t rlo[3], rhi[3];    // part of it inside this brick
    uint32_t key;              // bbType<<30 | triangle
    int32_t bb, axis;
};
constexpr int kSurfChunk = 64;

__device__ __forceinline__ bool test_plane(const OverlapData &o, int plane, const float pc[3]) {
    return plane == 0 ? overlap2(o, pc[1], pc[2])
         : plane == 1 ? overlap2(o, pc[2], pc[0])
                      : overlap2(o, pc[0], pc[1]);
}

// Sets up triangle t for the brick; returns its number of work items in the
// brick (voxels along the axis for 1D, dominant-plane columns for 2D/3D).
__device__ uint32_t surf_tri_setup(const SurfParams &p, uint32_t t, uint32_t info,
                                   const BrickBox &bx, SurfTri &st) {
    F3 tri[3];
    load_tri_f(p, t, tri);
    VoxBB vb;
    if (!surf_range(p, tri, vb)) return 0u;
    st.bb = (int32_t)(info >> 30);
    st.axis = (int32_t)((info >> 28) & 3u);
    st.key = ((uint32_t)st.bb << 30) | t;
    st.vmin[0] = vb.min.x; st.vmin[1] = vb.min.y; st.vmin[2] = vb.min.z;
    st.vmax[0] = vb.max.x; st.vmax[1] = vb.max.y; st.vmax[2] = vb.max.z;
    for (int a = 0; a < 3; a++) {
        st.rlo[a] = max(st.vmin[a], bx.lo[a]);
        st.rhi[a] = min(st.vmax[a], bx.hi[a]);
        if (st.rlo[a] > st.rhi[a]) return 0u;
    }
    const int ax = st.axis;
    if (st.bb == BB_1D) return (uint32_t)(st.rhi[ax] - st.rlo[ax] + 1);
    st.n = fcross(fsub(tri[0], tri[2]), fsub(tri[1], tri[0]));
    st.t0 = tri[0];
    setup_overlap(tri, st.n, p.d, 0, st.o[0]);
    setup_overlap(tri, st.n, p.d, 1, st.o[1]);
    setup_overlap(tri, st.n, p.d, 2, st.o[2]);
    const int ua = ax == AX_X ? AX_Y : AX_X;
    const int va = ax == AX_Z ? AX_Y : AX_Z;
    return (uint32_t)((st.rhi[ua] - st.rlo[ua] + 1) * (st.rhi[va] - st.rlo[va] + 1));
}

// One work item: per the 1D/2D/3D kernels of the reference
// (src/voxelizer.cu:3307-3955) with their matSplitRes subspace clipping --
// only the 3D single-voxel shortcut depends on the subspace, through the
// dominant-axis interval the voxel falls in.
__device__ void surf_item(const SurfParams &p, const SplitTable &spl, const SurfTri &st,
                          uint32_t k, const BrickBox &bx, uint32_t *keys) {
    const int ax = st.axis;
    if (st.bb == BB_1D) {
        int32_t c[3] = {st.vmin[0], st.vmin[1], st.vmin[2]};
        c[ax] = st.rlo[ax] + (int32_t)k;
        lds_write(keys, bx, p.bdim, c[0], c[1], c[2], st.key);
        return;
    }
    const int ua = ax == AX_X ? AX_Y : AX_X;
    const int va = ax == AX_Z ? AX_Y : AX_Z;
    const int32_t wu = st.rhi[ua] - st.rlo[ua] + 1;
    const int32_t u = st.rlo[ua] + (int32_t)k % wu;
    const int32_t v = st.rlo[va] + (int32_t)k / wu;
    const float mc[3] = {p.m.x, p.m.y, p.m.z};
    float pc[3];
    pc[ax] = 0.0f;
    pc[ua] = vox_corner(u, mc[ua], p.d);
    pc[va] = vox_corner(v, mc[va], p.d);
    if (!test_plane(st.o[ax], ax, pc)) return;
    int32_t c[3];
    c[ua] = u;
    c[va] = v;
    if (st.bb == BB_2D) {
        c[ax] = st.vmin[ax];
        lds_write(keys, bx, p.bdim, c[0], c[1], c[2], st.key);
        return;
    }
    int32_t lo, hi;
    depth_range(st.t0, F3{pc[0], pc[1], pc[2]}, st.n, p.m, p.d, ax, lo, hi);
    const int pa = ax == AX_X ? 1 : 0;
    const int pb = ax == AX_Z ? 1 : 2;
    const int32_t ilo = max(lo, st.rlo[ax]), ihi = min(hi, st.rhi[ax]);
    for (int32_t i = ilo; i <= ihi; i++) {
        int32_t slo;
        const int32_t shi = split_hi_of(spl, ax, i, slo);
        const int32_t bl = max(st.vmin[ax], slo), bh = min(st.vmax[ax], shi - 1);
        const int32_t rl = lo < bl ? bl : lo;
        const int32_t rh = hi > bh ? bh : hi;
        bool write = rl == rh;
        if (!write) {
            pc[ax] = vox_corner(i, mc[ax], p.d);
            write = test_plane(st.o[pa], pa, pc) && test_plane(st.o[pb], pb, pc);
        }
        if (write) {
            c[ax] = i;
            lds_write(keys, bx, p.bdim, c[0], c[1], c[2], st.key);
        }
    }
}

// One workgroup per brick.  The brick's triangles are taken 64 at a time:
// each is set up once into LDS, then all (triangle, column) work items of the
// chunk are spread over the 256 threads.  Winners accumulate in LDS
// (ds_max of bbType<<30 | triangle); the brick then writes its surface bit
// words, ORs them into the solid grid, and appends the materials of its
// surface voxels (compacted, voxel order) to the material list.
__global__ __launch_bounds__(256) void k_surf_brick(
    SurfParams p, const SplitTable *__restrict__ splits, const uint4 *__restrict__ tri_info,
    const uint32_t *__restrict__ brick_off, const uint32_t *__restrict__ brick_hist,
    const uint32_t *__restrict__ blist, const uint8_t *__restrict__ materials,
    uint32_t *__restrict__ bits, uint32_t *__restrict__ surf, uint32_t *__restrict__ brick_mat_off,
    uint8_t *__restrict__ mat_list, uint32_t *__restrict__ mat_cursor, uint32_t mat_cap,
    uint32_t *__restrict__ stats) {
    __shared__ uint32_t keys[kMaxBrickVox];
    __shared__ SurfTri tris[kSurfChunk];
    __shared__ uint32_t colsum[kSurfChunk + 1];
    __shared__ uint32_t wsurf[kMaxBrickVox / 32];
    __shared__ uint32_t wexcl[kMaxBrickVox / 32];
    __shared__ uint32_t sbase;
    __shared__ SplitTable spl;
    const uint32_t b = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t cnt = brick_hist[b];
    if (cnt == 0u) {
        if (tid == 0) brick_mat_off[b] = UINT32_MAX;
        return;
    }
    const uint32_t bxi = b % p.nb[0], byi = (b / p.nb[0]) % p.nb[1], bzi = b / (p.nb[0] * p.nb[1]);
    const int32_t nvox = p.bdim[0] * p.bdim[1] * p.bdim[2];
    for (int32_t i = tid; i < nvox; i += blockDim.x) keys[i] = 0u;
    {
        const uint32_t *src = (const uint32_t *)splits;
        uint32_t *dst = (uint32_t *)&spl;
        for (uint32_t i = tid; i < sizeof(SplitTable) / 4; i += blockDim.x) dst[i] = src[i];
    }
    BrickBox box;
    box.lo[0] = (int32_t)(bxi * p.bdim[0]) - 1;
    box.lo[1] = (int32_t)(byi * p.bdim[1]) - 1;
    box.lo[2] = (int32_t)(bzi * p.bdim[2]) - 1;
    for (int a = 0; a < 3; a++) box.hi[a] = box.lo[a] + p.bdim[a] - 1;
    const uint32_t off = brick_off[b];
    for (uint32_t c0 = 0; c0 < cnt; c0 += kSurfChunk) {
        __syncthreads();
        if (tid < (uint32_t)kSurfChunk) {
            uint32_t items = 0u;
            if (c0 + tid < cnt) {
                const uint32_t t = blist[off + c0 + tid];
                items = surf_tri_setup(p, t, tri_info[t].x, box, tris[tid]);
            }
            // wave-level inclusive scan of the item counts
            uint32_t incl = items;
            for (int d = 1; d < kSurfChunk; d <<= 1) {
                uint32_t o = (uint32_t)__shfl_up((int)incl, d);
                if ((int)tid >= d) incl += o;
            }
            colsum[tid + 1] = incl;
            if (tid == 0) colsum[0] = 0u;
        }
        __syncthreads();
        const uint32_t total = colsum[kSurfChunk];
        for (uint32_t c = tid; c < total; c += blockDim.x) {
            uint32_t lo = 0, hi = kSurfChunk;  // find i with colsum[i] <= c < colsum[i+1]
            while (lo + 1 < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (colsum[mid] <= c) lo = mid;
                else hi = mid;
            }
            surf_item(p, spl, tris[lo], c - colsum[lo], box, keys);
        }
    }
    __syncthreads();
    // Surface bit words: one ballot per 64 voxels (MSB-first words).
    const int32_t nwords = nvox / 32;
    for (int32_t v0 = (int32_t)(tid & ~63u); v0 < nvox; v0 += (int32_t)blockDim.x) {
        const unsigned long long m = __ballot(keys[v0 + (int32_t)lane_id()] != 0u);
        if (lane_id() == 0) {
            wsurf[v0 / 32] = __brev((uint32_t)m);
            wsurf[v0 / 32 + 1] = __brev((uint32_t)(m >> 32));
        }
    }
    __syncthreads();
    // Exclusive popcount scan over the brick's words.
    uint32_t pc = tid < (uint32_t)nwords ? (uint32_t)__popc(wsurf[tid]) : 0u;
    if (tid < (uint32_t)nwords) wexcl[tid] = pc;
    __syncthreads();
    for (uint32_t d = 1; d < (uint32_t)nwords; d <<= 1) {
        uint32_t o = (tid < (uint32_t)nwords && tid >= d) ? wexcl[tid - d] : 0u;
        __syncthreads();
        if (tid < (uint32_t)nwords) wexcl[tid] += o;
        __syncthreads();
    }
    if (tid == (uint32_t)nwords - 1) {
        const uint32_t acc = wexcl[tid];
        uint32_t base = atomicAdd(mat_cursor, acc);
        if (base + acc > mat_cap) {
            atomicOr(&stats[0], 1u);  // overflow: the caller retries with more room
            base = UINT32_MAX;
        }
        sbase = base;
        brick_mat_off[b] = base;
        atomicAdd(&stats[1], acc);
    }
    __syncthreads();
    const uint32_t base = sbase;
    const int32_t wpr = p.bdim[0] / 32;  // words per brick row
    for (int32_t w = tid; w < nwords; w += blockDim.x) {
        const int32_t wx = w % wpr, ly = (w / wpr) % p.bdim[1], lz = w / (wpr * p.bdim[1]);
        const uint32_t aw = bxi * (uint32_t)wpr + (uint32_t)wx;
        const uint32_t ay = byi * (uint32_t)p.bdim[1] + (uint32_t)ly;
        const uint32_t az = bzi * (uint32_t)p.bdim[2] + (uint32_t)lz;
        if (aw >= p.alloc[0] / 32u || ay >= p.alloc[1] || az >= p.alloc[2]) continue;
        const size_t gi = ((size_t)az * p.alloc[1] + ay) * (p.alloc[0] / 32u) + aw;
        const uint32_t s = wsurf[w];
        surf[gi] = s;
        if (s) bits[gi] = bits[gi] | s;
        if (base != UINT32_MAX && s) {
            uint32_t r = base + wexcl[w] - (uint32_t)__popc(s);
            uint32_t m = s;
            while (m) {
                const int k = __clz(m);
                const uint32_t key = keys[w * 32 + k];
                mat_list[r++] = materials ? materials[key & 0x3FFFFFFFu] : (uint8_t)0;
                m &= ~(0x80000000u >> k);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Erosion: one snapshot pass of "solid and a solid neighbour on every axis"
// over bit words (the invalid-permutation removal of src/voxelizer.cu:
// 1791-1913, iterated to its fixed point by the host as in
// src/host_side.cpp:1183-1206).  Words outside [y0,y1) x [z0,z1) are copied.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ldw(const uint32_t *__restrict__ a, const GridDims &g,
                                        int64_t w, int64_t y, int64_t z) {
    if (w < 0 || w >= (int64_t)g.W || y < 0 || y >= (int64_t)g.Y || z < 0 || z >= (int64_t)g.Z)
        return 0u;
    return a[((size_t)z * g.Y + (size_t)y) * g.W + (size_t)w];
}

__global__ __launch_bounds__(256) void k_erode(GridDims g, const uint32_t *__restrict__ src,
                                               uint32_t *__restrict__ dst, uint32_t y0,
                                               uint32_t y1, uint32_t z0, uint32_t z1,
                                               uint32_t *__restrict__ changes, uint32_t pass) {
    if (pass > 0 && changes[pass - 1] == 0u) return;
    const size_t total = (size_t)g.W * g.Y * g.Z;
    for (size_t base = (size_t)blockIdx.x * blockDim.x; base < total;
         base += (size_t)gridDim.x * blockDim.x) {
        const size_t i = base + threadIdx.x;
        bool changed = false;
        if (i < total) {
            const int64_t w = (int64_t)(i % g.W);
            const int64_t y = (int64_t)((i / g.W) % g.Y);
            const int64_t z = (int64_t)(i / ((size_t)g.W * g.Y));
            const uint32_t G = src[i];
            uint32_t out = G;
            if (G && (uint64_t)y >= y0 && (uint64_t)y < y1 && (uint64_t)z >= z0 &&
                (uint64_t)z < z1) {
                const uint32_t L = (G >> 1) | (ldw(src, g, w - 1, y, z) << 31);
                const uint32_t R = (G << 1) | (ldw(src, g, w + 1, y, z) >> 31);
                const uint32_t Yn = ldw(src, g, w, y - 1, z) | ldw(src, g, w, y + 1, z);
                const uint32_t Zn = ldw(src, g, w, y, z - 1) | ldw(src, g, w, y, z + 1);
                out = G & (L | R) & Yn & Zn;
                changed = out != G;
            }
            dst[i] = out;
        }
        const unsigned long long m = __ballot(changed);
        if (lane_id() == 0 && m) atomicAdd(&changes[pass], (uint32_t)__popcll(m));
    }
}

// ---------------------------------------------------------------------------
// Node writing
// ---------------------------------------------------------------------------
__constant__ uint8_t c_orient[64];

hipError_t set_orient_table(const uint8_t lut[64]) {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_orient), lut, 64);
}

// Rank of each surface voxel inside its brick: exclusive popcount prefix of
// the brick's surface words (same word order as k_surf_brick).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t n) {
    const uint32_t tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < n; off <<= 1) {
        uint32_t t = (tid >= off && tid < n) ? sh[tid - off] : 0u;
        __syncthreads();
        if (tid < n) sh[tid] += t;
        __syncthreads();
    }
    uint32_t r = (tid > 0 && tid < n) ? sh[tid - 1] : 0u;
    __syncthreads();
    return r;
}

// 4 bits (MSB = first node) -> 4 bytes of 0/1 in little-endian node order.
__device__ __forceinline__ uint32_t spread4(uint32_t nib) {
    return ((nib >> 3) & 1u) | (((nib >> 2) & 1u) << 8) | (((nib >> 1) & 1u) << 16) |
           ((nib & 1u) << 24);
}
// 2 bits -> 2 halfwords of 0/1.
__device__ __forceinline__ uint32_t spread2(uint32_t v) {
    return ((v >> 1) & 1u) | ((v & 1u) << 16);
}

// One 128-thread workgroup per 128x4x8 brick, one thread per 32-node bit
// word: six neighbour masks from bit shifts of the word and its five
// neighbour words, boundary ids (27 for fully surrounded nodes, the 64-entry
// table for the rest), materials from the brick's surface list, and one
// vector store of the 32 packed nodes.
//
// With several erosion subspaces (voxSplitRes y/z smaller than the grid) the
// reference computes a subspace's ids before later subspaces erode
// (src/host_side.cpp:1183-1206), so neighbours across a +y/+z seam are read
// from the pre-erosion snapshot `pre` (seam rows flagged in ystart/zstart).
__global__ __launch_bounds__(128) void k_nodes_cart(GridDims g, const uint32_t *__restrict__ solid,
                                                    const uint32_t *__restrict__ pre,
                                                    const uint8_t *__restrict__ ystart,
                                                    const uint8_t *__restrict__ zstart,
                                                    MatSource ms, int use_mats, int kind,
                                                    void *__restrict__ nodes, uint32_t nbx,
                                                    uint32_t nby) {
    __shared__ uint32_t scan_sh[128];
    const uint32_t tid = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    const uint32_t wx = tid & 3u, ly = (tid >> 2) & 3u, lz = tid >> 4;
    const uint32_t aw = bxi * 4u + wx, ay = byi * kCartBY + ly, az = bzi * kCartBZ + lz;
    const bool valid = aw < g.W && ay < g.Y && az < g.Z;
    uint32_t G = 0, L = 0, R = 0, Yn = 0, Yp = 0, Zn = 0, Zp = 0;
    const size_t wi = ((size_t)az * g.Y + ay) * g.W + aw;
    if (valid) {
        G = solid[wi];
        L = (G >> 1) | (ldw(solid, g, (int64_t)aw - 1, ay, az) << 31);
        R = (G << 1) | (ldw(solid, g, (int64_t)aw + 1, ay, az) >> 31);
        Yn = ldw(solid, g, aw, (int64_t)ay - 1, az);
        const bool ys = pre && ay + 1 < g.Y && ystart[ay + 1];
        const bool zs = pre && az + 1 < g.Z && zstart[az + 1];
        Yp = ldw(ys ? pre : solid, g, aw, (int64_t)ay + 1, az);
        Zn = ldw(solid, g, aw, ay, (int64_t)az - 1);
        Zp = ldw(zs ? pre : solid, g, aw, ay, (int64_t)az + 1);
    }
    // Materials.
    uint32_t S = 0u, mrank = 0u, mbase = UINT32_MAX;
    if (use_mats) {
        mbase = ms.brick_mat_off[b];
        if (mbase != UINT32_MAX) {
            S = valid ? ms.surf[wi] : 0u;
            mrank = block_excl_scan((uint32_t)__popc(S), scan_sh, 128u);
        }
    }
    if (!valid) return;
    const uint32_t interior = G & L & R & Yn & Yp & Zn & Zp;
    uint32_t boundary = G & ~interior;
    const size_t node0 = wi * 32u;
    if (kind == 0) {  // ShortNode: mat<<5 | bid
        uint32_t out[8];
#pragma unroll
        for (int j = 0; j < 8; j++) out[j] = spread4((interior >> (28 - 4 * j)) & 15u) * 27u;
        while (boundary) {
            const int k = __clz(boundary);
            const uint32_t bit = 0x80000000u >> k;
            const uint32_t perm = ((L & bit) ? 1u : 0u) | ((R & bit) ? 2u : 0u) |
                                  ((Yn & bit) ? 4u : 0u) | ((Yp & bit) ? 8u : 0u) |
                                  ((Zn & bit) ? 16u : 0u) | ((Zp & bit) ? 32u : 0u);
            out[k >> 2] |= (uint32_t)c_orient[perm] << (8 * (k & 3));
            boundary &= ~bit;
        }
        if (S) {
            uint32_t r = mbase + mrank;
            uint32_t m = S;
            while (m) {
                const int k = __clz(m);
                const uint32_t mat = ms.mat_list[r++];
                out[k >> 2] |= ((mat << 5) & 0xFFu) << (8 * (k & 3));
                m &= ~(0x80000000u >> k);
            }
        }
        uint4 *dst = (uint4 *)((uint8_t *)nodes + node0);
        dst[0] = make_uint4(out[0], out[1], out[2], out[3]);
        dst[1] = make_uint4(out[4], out[5], out[6], out[7]);
    } else {  // LongNode: {bid, mat}
        uint32_t out[16];
#pragma unroll
        for (int j = 0; j < 16; j++) out[j] = spread2((interior >> (30 - 2 * j)) & 3u) * 27u;
        while (boundary) {
            const int k = __clz(boundary);
            const uint32_t bit = 0x80000000u >> k;
            const uint32_t perm = ((L & bit) ? 1u : 0u) | ((R & bit) ? 2u : 0u) |
                                  ((Yn & bit) ? 4u : 0u) | ((Yp & bit) ? 8u : 0u) |
                                  ((Zn & bit) ? 16u : 0u) | ((Zp & bit) ? 32u : 0u);
            out[k >> 1] |= (uint32_t)c_orient[perm] << (16 * (k & 1));
            boundary &= ~bit;
        }
        if (S) {
            uint32_t r = mbase + mrank;
            uint32_t m = S;
            while (m) {
                const int k = __clz(m);
                const uint32_t mat = ms.mat_list[r++];
                out[k >> 1] |= mat << (16 * (k & 1) + 8);
                m &= ~(0x80000000u >> k);
            }
        }
        uint4 *dst = (uint4 *)((uint8_t *)nodes + node0 * 2u);
#pragma unroll
        for (int j = 0; j < 4; j++)
            dst[j] = make_uint4(out[4 * j], out[4 * j + 1], out[4 * j + 2], out[4 * j + 3]);
    }
}


// ---------------------------------------------------------------------------
// FCC nodes
// ---------------------------------------------------------------------------
// Spreads 16 bits to the odd... even positions: bit j -> bit 2j.
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

struct FccGrids {
    const uint32_t *s[4];  // per-grid solid|surface bits (grid 1..4)
};

// 32 FCC nodes of row (y, fz), x in [32*fw, 32*fw+32): even x from the
// sub-grid with x shift 0, odd x from the one with x shift 1/2
// (convertToFCCGrid, src/voxelizer.cu:1756-1772).
__device__ __forceinline__ uint32_t fcc_word(const FccGrids &G, const GridDims &g, int64_t fw,
                                             int64_t y, int64_t fz) {
    if (fw < 0 || fw >= (int64_t)(2 * g.W) || y < 0 || y >= (int64_t)g.Y || fz < 0 ||
        fz >= (int64_t)(2 * g.Z))
        return 0u;
    const int odd = (int)(fz & 1);
    const uint32_t *ge = G.s[odd ? 2 : 0], *go = G.s[odd ? 3 : 1];
    const size_t wi = ((size_t)(fz >> 1) * g.Y + (size_t)y) * g.W + (size_t)(fw >> 1);
    const int sh = (fw & 1) ? 0 : 16;
    const uint32_t e = (ge[wi] >> sh) & 0xFFFFu, o = (go[wi] >> sh) & 0xFFFFu;
    return (spread16(e) << 1) | spread16(o);
}

// One 128-thread workgroup per 128x4x8 FCC brick (= one 64x4x4 brick of each
// sub-grid), one thread per 32-node word: the 12-bit neighbour mask of
// calculateFCCBoundaries (src/voxelizer.cu:1930-2048) as bitwise selects by
// the (x+z) parity mask, materials from the four sub-grids' surface lists.
__global__ __launch_bounds__(128) void k_nodes_fcc(GridDims g, const uint32_t *s1,
                                                   const uint32_t *s2, const uint32_t *s3,
                                                   const uint32_t *s4, MatSource m1, MatSource m2,
                                                   MatSource m3, MatSource m4, int use_mats,
                                                   int kind, void *__restrict__ nodes,
                                                   uint32_t nbx, uint32_t nby) {
    __shared__ uint32_t rank_sh[128];
    __shared__ uint32_t mbase_sh[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint32_t bxi = b % nbx, byi = (b / nbx) % nby, bzi = b / (nbx * nby);
    const uint32_t fwl = tid & 3u, ly = (tid >> 2) & 3u, lz = tid >> 4;
    const int64_t fw = (int64_t)bxi * 4 + fwl;
    const int64_t y = (int64_t)byi * 4 + ly;
    const int64_t fz = (int64_t)bzi * 8 + lz;
    const uint32_t FW = 2 * g.W, FY = g.Y, FZ = 2 * g.Z;
    const bool valid = fw < (int64_t)FW && y < (int64_t)FY && fz < (int64_t)FZ;
    FccGrids G{{s1, s2, s3, s4}};
    const MatSource ms[4] = {m1, m2, m3, m4};

    // Surface ranks: thread t handles word (t & 31) of sub-grid (t >> 5)'s
    // brick (64x4x4 voxels, word order (lz*4 + ly)*2 + wx).
    uint32_t gsurf = 0u;
    if (use_mats) {
        const uint32_t gi = tid >> 5, w = tid & 31u;
        const uint32_t wx = w & 1u, gly = (w >> 1) & 3u, glz = w >> 3;
        const uint64_t aw = (uint64_t)bxi * 2 + wx, ay = (uint64_t)byi * 4 + gly,
                       az = (uint64_t)bzi * 4 + glz;
        const uint32_t mb = ms[gi].brick_mat_off[b];
        if (tid < 4) mbase_sh[tid] = ms[tid].brick_mat_off[b];
        if (mb != UINT32_MAX && aw < g.W && ay < g.Y && az < g.Z)
            gsurf = ms[gi].surf[(az * g.Y + ay) * g.W + aw];
        rank_sh[tid] = (uint32_t)__popc(gsurf);
        __syncthreads();
        // Segmented (per sub-grid, 32 entries) inclusive scan.
        for (uint32_t off = 1; off < 32; off <<= 1) {
            uint32_t t = (w >= off) ? rank_sh[tid - off] : 0u;
            __syncthreads();
            rank_sh[tid] += t;
            __syncthreads();
        }
        rank_sh[tid] -= (uint32_t)__popc(gsurf);  // exclusive
        __syncthreads();
    }
    if (!valid) return;

    const uint32_t M0 = fcc_word(G, g, fw, y, fz);
    const uint32_t Mm = fcc_word(G, g, fw, y - 1, fz), Mp = fcc_word(G, g, fw, y + 1, fz);
    const uint32_t D0 = fcc_word(G, g, fw, y, fz - 1);
    const uint32_t Dm = fcc_word(G, g, fw, y - 1, fz - 1), Dp = fcc_word(G, g, fw, y + 1, fz - 1);
    const uint32_t U0 = fcc_word(G, g, fw, y, fz + 1);
    const uint32_t Um = fcc_word(G, g, fw, y - 1, fz + 1), Up = fcc_word(G, g, fw, y + 1, fz + 1);
    auto Lsh = [&](uint32_t w, int64_t yy, int64_t zz) {
        return (w >> 1) | (fcc_word(G, g, fw - 1, yy, zz) << 31);
    };
    auto Rsh = [&](uint32_t w, int64_t yy, int64_t zz) {
        return (w << 1) | (fcc_word(G, g, fw + 1, yy, zz) >> 31);
    };
    const uint32_t E = (fz & 1) ? 0x55555555u : 0xAAAAAAAAu;  // (x+z) even
    const uint32_t NE = ~E;
    uint32_t n[12];
    n[0] = (E & Dm) | (NE & D0);
    n[1] = Lsh(D0, y, fz - 1);
    n[2] = Rsh(D0, y, fz - 1);
    n[3] = (E & D0) | (NE & Dp);
    const uint32_t LMm = Lsh(Mm, y - 1, fz), LM0 = Lsh(M0, y, fz), LMp = Lsh(Mp, y + 1, fz);
    const uint32_t RMm = Rsh(Mm, y - 1, fz), RM0 = Rsh(M0, y, fz), RMp = Rsh(Mp, y + 1, fz);
    n[4] = (E & LMm) | (NE & LM0);
    n[5] = (E & RMm) | (NE & RM0);
    n[6] = (E & LM0) | (NE & LMp);
    n[7] = (E & RM0) | (NE & RMp);
    n[8] = (E & Um) | (NE & U0);
    n[9] = Lsh(U0, y, fz + 1);
    n[10] = Rsh(U0, y, fz + 1);
    n[11] = (E & U0) | (NE & Up);
    uint32_t interior = M0;
#pragma unroll
    for (int j = 0; j < 12; j++) interior &= n[j];
    uint32_t mask[32];
#pragma unroll
    for (int k = 0; k < 32; k++) mask[k] = ((interior >> (31 - k)) & 1u) ? 4095u : 0u;
    uint32_t boundary = M0 & ~interior;
    while (boundary) {
        const int k = __clz(boundary);
        const uint32_t bit = 0x80000000u >> k;
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 12; j++) m |= ((n[j] & bit) ? 1u : 0u) << j;
#pragma unroll
        for (int kk = 0; kk < 32; kk++)
            if (kk == k) mask[kk] = m;
        boundary &= ~bit;
    }
    uint32_t mat[32];
#pragma unroll
    for (int k = 0; k < 32; k++) mat[k] = 0u;
    if (use_mats) {
        const int odd = (int)(fz & 1);
        const int gsel[2] = {odd ? 2 : 0, odd ? 3 : 1};  // even x, odd x
        const uint32_t glz = (uint32_t)(lz >> 1);
        const uint32_t wx = (uint32_t)(fwl >> 1);
        const uint32_t gw = (glz * 4 + ly) * 2 + wx;
        const uint64_t aw = (uint64_t)bxi * 2 + wx, az = (uint64_t)(fz >> 1);
        const int sh = (fwl & 1) ? 0 : 16;
        for (int par = 0; par < 2; par++) {
            const int gi = gsel[par];
            const uint32_t mb = mbase_sh[gi];
            if (mb == UINT32_MAX) continue;
            const uint32_t sw = ms[gi].surf[(az * g.Y + (uint64_t)y) * g.W + aw];
            // voxels of this half, MSB first; rank counts set bits before it
            uint32_t r = mb + rank_sh[gi * 32 + gw] + ((fwl & 1) ? (uint32_t)__popc(sw >> 16) : 0u);
            uint32_t h = (sw >> sh) & 0xFFFFu;
            while (h) {
                const int i = __clz(h) - 16;  // voxel i of the 16
                const uint32_t mv = ms[gi].mat_list[r++];
                const int k = 2 * i + par;
#pragma unroll
                for (int kk = 0; kk < 32; kk++)
                    if (kk == k) mat[kk] = mv;
                h &= ~(0x8000u >> i);
            }
        }
    }
    const size_t node0 = (((size_t)fz * FY + (size_t)y) * FW + (size_t)fw) * 32u;
    if (kind == 3) {  // ShortFCCNode: mat<<12 | bid&4095
        uint32_t out[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            uint32_t lo = ((mat[2 * j] << 12) | (mask[2 * j] & 4095u)) & 0xFFFFu;
            uint32_t hi = ((mat[2 * j + 1] << 12) | (mask[2 * j + 1] & 4095u)) & 0xFFFFu;
            out[j] = lo | (hi << 16);
        }
        uint4 *dst = (uint4 *)((uint8_t *)nodes + node0 * 2u);
#pragma unroll
        for (int j = 0; j < 4; j++)
            dst[j] = make_uint4(out[4 * j], out[4 * j + 1], out[4 * j + 2], out[4 * j + 3]);
    } else {  // LongFCCNode: {u16 bid, u8 mat, u8 0}
        uint4 *dst = (uint4 *)((uint8_t *)nodes + node0 * 4u);
#pragma unroll
        for (int j = 0; j < 8; j++)
            dst[j] = make_uint4(mask[4 * j] | (mat[4 * j] & 0xFFu) << 16,
                                mask[4 * j + 1] | (mat[4 * j + 1] & 0xFFu) << 16,
                                mask[4 * j + 2] | (mat[4 * j + 2] & 0xFFu) << 16,
                                mask[4 * j + 3] | (mat[4 * j + 3] & 0xFFu) << 16);
    }
}

}  // namespace vk
