// vox_kernels.hip -- HIP kernels of the voxelizer hot path for gfx950.
// The arithmetic of every test follows the reference expressions cited in
// vox_math.h; the data flow (binning, LDS staging, bitwise node stencils) is
// this library's own.
#include "vox_kernels.h"

#include <type_traits>

namespace vk {

// ---------------------------------------------------------------------------
// Scan
// ---------------------------------------------------------------------------
// Every scan writes out[0, n] (exclusive prefix, out[n] = total mod 2^32) and
// out[n + 1] = 1 if the true total reached 2^32 (the 32-bit offsets wrapped),
// else 0.  Tile sums are kept in 64 bits for that: partials[b] holds tile b's
// low word and partials[ntiles + b] its high word.

// Sum of one value per thread over a 256-thread block, in 64 bits.
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t s) {
    __shared__ uint64_t wsum64[kScanThreads / kWave];
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)s, d);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(s >> 32), d);
        s += ((uint64_t)hi << 32) | lo;
    }
    if ((threadIdx.x & (kWave - 1u)) == 0u) wsum64[threadIdx.x / kWave] = s;
    __syncthreads();
    uint64_t t = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanThreads / kWave; w++) t += wsum64[w];
    __syncthreads();  // wsum64 may be reused by the caller's next call
    return t;
}

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total); one block scans
// kScanTile elements and leaves its sum in partials[block] (low word) and
// partials[nblocks + block] (high word).  The last block clears the wrap flag
// out[n + 1]; k_scan_add sets it.
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const uint32_t *__restrict__ in,
                                                              uint32_t *__restrict__ out,
                                                              uint32_t n,
                                                              uint32_t *__restrict__ partials) {
    __shared__ uint32_t sh[kScanThreads];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)tid * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    uint64_t s64 = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        uint64_t idx = base + k;
        v[k] = idx < n ? in[idx] : 0u;
        s += v[k];
        s64 += v[k];
    }
    const uint64_t tile64 = block_sum_u64(s64);
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
    if (tid == kScanThreads - 1) {
        partials[blockIdx.x] = sh[tid];
        partials[gridDim.x + blockIdx.x] = (uint32_t)(tile64 >> 32);
    }
    if (tid == 0 && blockIdx.x == gridDim.x - 1u) out[(uint64_t)n + 1u] = 0u;
}

// Block-wide sum and exclusive prefix of one value per thread: a wave scan
// in registers, the four wave totals through LDS (one barrier).
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t s, uint32_t &total) {
    __shared__ uint32_t wsum[kScanThreads / kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1u), wave = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(s);
    if (lane == kWave - 1u) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < kScanThreads / kWave; w++) {
        before += w < wave ? wsum[w] : 0u;
        total += wsum[w];
    }
    return before + incl - s;
}

// Tile sums of in[0, n) (direct path): low words in partials[0, nblocks),
// high words in partials[nblocks, 2 * nblocks).
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                              uint32_t n,
                                                              uint32_t *__restrict__ partials) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t s = 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t idx = base + k;
        s += idx < n ? in[idx] : 0u;
    }
    const uint64_t total = block_sum_u64(s);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = (uint32_t)total;
        partials[gridDim.x + blockIdx.x] = (uint32_t)(total >> 32);
    }
}

// Exclusive scan of in[0, n) into out[0, n] (out[n] = total), direct path:
// the tile first sums the tile sums before it (at most kScanDirectTiles).
// The last tile also adds them in 64 bits and writes the wrap flag.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t *__restrict__ in,
                                                            uint32_t *__restrict__ out,
                                                            uint32_t n,
                                                            const uint32_t *__restrict__ partials) {
    __shared__ uint32_t s_off;
    const bool last = blockIdx.x == gridDim.x - 1u;
    uint32_t p = 0u;
    uint64_t p64 = 0u;
    for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kScanThreads) {
        p += partials[i];
        if (last) p64 += ((uint64_t)partials[gridDim.x + i] << 32) | partials[i];
    }
    uint32_t ptotal;
    block_excl_scan(p, ptotal);
    if (threadIdx.x == 0) s_off = ptotal;
    __syncthreads();  // also orders the two uses of block_excl_scan's LDS
    const uint32_t off = s_off;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t idx = base + k;
        v[k] = idx < n ? in[idx] : 0u;
        s += v[k];
        if (last) p64 += v[k];
    }
    uint32_t total;
    uint32_t excl = off + block_excl_scan(s, total);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const uint64_t idx = base + k;
        if (idx <= n) out[idx] = excl;
        excl += v[k];
    }
    if (last) {  // uniform
        const uint64_t all = block_sum_u64(p64);
        if (threadIdx.x == 0) out[(uint64_t)n + 1u] = (all >> 32) != 0u ? 1u : 0u;
    }
}

// Adds the scanned tile sums (recursive path) and sets the wrap flag when a
// tile's sum or the sum of the tile sums reached 2^32 (psum[nblocks + 1] is
// the recursive scan's own flag).
__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t *__restrict__ out,
                                                           uint32_t n,
                                                           const uint32_t *__restrict__ partials,
                                                           const uint32_t *__restrict__ psum) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    const uint32_t add = psum[blockIdx.x];
    for (uint32_t k = threadIdx.x; k < (uint32_t)kScanTile; k += kScanThreads) {
        uint64_t idx = base + k;
        if (idx <= n) out[idx] += add;
    }
    if (threadIdx.x == 0 &&
        (partials[gridDim.x + blockIdx.x] != 0u || (blockIdx.x == 0 && psum[gridDim.x + 1u] != 0u)))
        out[(uint64_t)n + 1u] = 1u;
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
__global__ __launch_bounds__(256) void k_tri_records(const Job *__restrict__ jobs,
                                                     const uint32_t *__restrict__ tri_bases,
                                                     uint32_t njobs, uint32_t nslots,
                                                     float4 *__restrict__ rec,
                                                     uint32_t *__restrict__ tri_job) {
    for (uint32_t t0 = blockIdx.x * 256u; t0 < nslots; t0 += gridDim.x * 256u) {
        const uint32_t j = job_of(tri_bases, njobs, t0);
        const Job &J = jobs[j];
        const uint32_t t = t0 + threadIdx.x, lt = t - J.tri_base;
        tri_job[t] = j;
        if (lt >= J.nt) continue;  // padding slot
        const VK_G float *V = gptr(J.V);
        const VK_G uint32_t *I = gptr(J.I);
        const uint32_t a = I[3u * lt], b = I[3u * lt + 1u], c = I[3u * lt + 2u];
        rec[3 * (size_t)t] = make_float4(V[3u * a], V[3u * a + 1u], V[3u * a + 2u], V[3u * b]);
        rec[3 * (size_t)t + 1] =
            make_float4(V[3u * b + 1u], V[3u * b + 2u], V[3u * c], V[3u * c + 1u]);
        const uint32_t mat = J.mats ? (uint32_t)gptr(J.mats)[lt] : 0u;
        rec[3 * (size_t)t + 2] = make_float4(V[3u * c + 2u], 0.0f, __uint_as_float(mat), 0.0f);
    }
}

// The triangle's vertices widened to FP64 (fetchTriangleVertices,
// src/voxelizer.cu:2242-2253), from its 48-byte record.
__device__ __forceinline__ void load_tri_d(const SolidParams &p, uint32_t t, D3 v[3]) {
    const float4 r0 = p.rec[3 * (size_t)t], r1 = p.rec[3 * (size_t)t + 1],
                 r2 = p.rec[3 * (size_t)t + 2];
    v[0] = {(double)r0.x, (double)r0.y, (double)r0.z};
    v[1] = {(double)r0.w, (double)r1.x, (double)r1.y};
    v[2] = {(double)r1.z, (double)r1.w, (double)r2.x};
}

// Candidate tiles of each triangle: its commonSpace tile rectangle; the
// commonSpace box itself is kept so the candidate pass needs no divisions.
__global__ __launch_bounds__(256) void k_solid_setup(const Job *__restrict__ jobs,
                                                     const uint32_t *__restrict__ tri_bases,
                                                     uint32_t njobs, uint32_t nslots,
                                                     const float4 *__restrict__ rec,
                                                     uint32_t *__restrict__ cand_count,
                                                     uint4 *__restrict__ cbox) {
    for (uint32_t t0 = blockIdx.x * 256u; t0 < nslots; t0 += gridDim.x * 256u) {
        const Job &J = jobs[job_of(tri_bases, njobs, t0)];
        const SolidParams p = solid_params(J, rec);
        const uint32_t t = t0 + threadIdx.x;
        uint32_t c = 0;
        if (t - J.tri_base < J.nt) {
            D3 v[3];
            load_tri_d(p, t, v);
            TileSetup ts;
            if (tile_setup(v, p.emy, p.emz, p.d, p.smin[0], p.smin[1], p.smax[0], p.smax[1],
                           ts)) {
                c = (ts.tmax[0] - ts.tmin[0] + 1u) * (ts.tmax[1] - ts.tmin[1] + 1u);
                cbox[t] = make_uint4(ts.cmin[0], ts.cmin[1], ts.cmax[0], ts.cmax[1]);
            }
        }
        cand_count[t] = c;
    }
}

// One thread per candidate (triangle, tile): the reference's per-tile queue
// test, done in parallel instead of one thread walking all tiles of a
// triangle.  The triangle of a candidate is found by a binary search narrowed
// to the triangles the block's candidates span.  Queued pairs are counted per
// tile (LDS-grouped atomics) for the counting sort.
// Block starts of a candidate pass: first[b] = the triangle that holds
// candidate 256*b (the passes' blocks take 256 consecutive candidates), so a
// block bounds its triangles with two loads instead of a binary search over
// the whole offset array (about 21 dependent loads on the large hall).
// Candidates at or past `cap` are not processed.
__global__ __launch_bounds__(256) void k_block_starts(const uint32_t *__restrict__ cand_off,
                                                      uint32_t nslots, uint32_t cap,
                                                      uint32_t *__restrict__ first) {
    const uint32_t total = cand_off[nslots + 1u] ? 0u : min(cand_off[nslots], cap);
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < nslots; t += gridDim.x * 256u) {
        const uint32_t lo = min(cand_off[t], total), hi = min(cand_off[t + 1u], total);
        for (uint32_t b = (lo + 255u) / 256u; b * 256u < hi; b++) first[b] = t;
    }
}

// Rows of tile (ty, tz) whose voxSplitRes subspace queues the triangle (bit
// r = centre (4 ty + (r & 3), 4 tz + (r >> 2))).  The reference builds its
// queue per y/z subspace (src/host_side.cpp:275-397): a tile is queued for a
// subspace when one of its centres inside that subspace passes the (>= 0)
// filter (src/voxelizer.cu:880-925, 998-1031), and the fill then tests only
// that subspace's centres.  So a centre's crossing counts only when some
// centre of the same subspace part of the tile passes.  sy/sz: the subspace
// start flags of the ext rows (Job::sstart).
__device__ __forceinline__ uint32_t tile_split_mask(const TileSetup &ts, uint32_t ty, uint32_t tz,
                                                    double emy, double emz, double d,
                                                    const uint8_t *__restrict__ sy, uint32_t ry,
                                                    const uint8_t *__restrict__ sz, uint32_t rz) {
    uint32_t pass = 0u;
#pragma unroll
    for (uint32_t r = 0; r < 16u; r++) {
        const uint32_t i = ty * 4u + (r & 3u), j = tz * 4u + (r >> 2);
        if (i >= ts.cmin[0] && i < ts.cmax[0] && j >= ts.cmin[1] && j < ts.cmax[1]) {
            const double px = double(i) * d + emy;
            const double py = double(j) * d + emz;
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 3; k++)
                ok &= (ts.ne[k].x * px + ts.ne[k].y * py + ts.de[k] >= 0.0);
            pass |= (ok ? 1u : 0u) << r;
        }
    }
    if (!pass) return 0u;
    // Subspace parts of the tile's 4 y rows and 4 z rows: part ids grow by
    // one where a subspace starts.
    uint32_t gy[4], gz[4];
    gy[0] = gz[0] = 0u;
#pragma unroll
    for (uint32_t k = 1; k < 4u; k++) {
        const uint32_t y = ty * 4u + k, z = tz * 4u + k;
        gy[k] = gy[k - 1] + (y < ry && sy[y] ? 1u : 0u);
        gz[k] = gz[k - 1] + (z < rz && sz[z] ? 1u : 0u);
    }
    uint32_t ysame[4], zsame[4];  // rows in the same part as row k (4-bit sets)
#pragma unroll
    for (uint32_t k = 0; k < 4u; k++) {
        ysame[k] = zsame[k] = 0u;
#pragma unroll
        for (uint32_t q = 0; q < 4u; q++) {
            ysame[k] |= (gy[q] == gy[k] ? 1u : 0u) << q;
            zsame[k] |= (gz[q] == gz[k] ? 1u : 0u) << q;
        }
    }
    uint32_t mask = 0u;
#pragma unroll
    for (uint32_t r = 0; r < 16u; r++) {
        if (!((pass >> r) & 1u)) continue;
        const uint32_t ys = ysame[r & 3u], zs = zsame[r >> 2];
        const uint32_t yrows = ys | (ys << 4) | (ys << 8) | (ys << 12);
        const uint32_t zrows = ((zs & 1u) ? 0x000Fu : 0u) | ((zs & 2u) ? 0x00F0u : 0u) |
                               ((zs & 4u) ? 0x0F00u : 0u) | ((zs & 8u) ? 0xF000u : 0u);
        mask |= yrows & zrows;
    }
    return mask;
}

template <bool kMulti>
__global__ __launch_bounds__(256) void k_solid_candidates(
    const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_job, uint32_t nslots,
    const float4 *__restrict__ rec, const uint32_t *__restrict__ cand_off,
    const uint32_t *__restrict__ cand_total, const uint32_t *__restrict__ first,
    const uint4 *__restrict__ cbox, uint2 *__restrict__ pairs, uint16_t *__restrict__ pmask,
    uint32_t *__restrict__ tile_hist, uint32_t cap, uint32_t *__restrict__ ovf) {
    __shared__ uint32_t trange[2];
    __shared__ AggTable agg;
    // cand_total[1]: the candidate count reached 2^32 (the offsets wrapped).
    const uint32_t full = *cand_total, total = cand_total[1] ? 0u : min(full, cap);
    if (ovf && (full > cap || cand_total[1]) && blockIdx.x == 0 && threadIdx.x == 0) *ovf = 1u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += stride) {
        __syncthreads();
        agg_clear(agg);
        if (threadIdx.x < 2) {  // the triangles of candidates base and base + 256 bound the block's
            const uint32_t b = base / 256u + threadIdx.x;
            trange[threadIdx.x] = threadIdx.x == 0 || base + 256u < total ? first[b] : nslots - 1u;
        }
        __syncthreads();
        const uint32_t c = base + threadIdx.x;
        bool hit = false;
        uint32_t t = 0, tile = 0;
        if (c < total) {
            uint32_t lo = trange[0], hi = trange[1] + 1u;
            while (lo + 1u < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (cand_off[mid] <= c) lo = mid;
                else hi = mid;
            }
            t = lo;
            const Job &J = jobs[kMulti ? tri_job[t] : 0u];
            const SolidParams p = solid_params(J, rec);
            D3 v[3];
            load_tri_d(p, t, v);
            TileSetup ts;
            const uint4 cb = cbox[t];
            ts.cmin[0] = cb.x;
            ts.cmin[1] = cb.y;
            ts.cmax[0] = cb.z;
            ts.cmax[1] = cb.w;
            tile_edges(v, ts);
            const uint32_t tmin0 = cb.x / 4u, tmin1 = cb.y / 4u, tmax0 = (cb.z - 1u) / 4u;
            const uint32_t w = tmax0 - tmin0 + 1u;
            const uint32_t local = c - cand_off[t];
            const uint32_t ty = tmin0 + local % w, tz = tmin1 + local / w;
            if (J.sstart) {  // several y/z subspaces
                const uint32_t m = tile_split_mask(ts, ty, tz, p.emy, p.emz, p.d, J.sstart,
                                                   (uint32_t)J.res[1], J.sstart + J.res[1],
                                                   (uint32_t)J.res[2]);
                hit = m != 0u;
                pmask[c] = (uint16_t)m;
            } else {
                hit = tile_hit(ts, ty, tz, p.emy, p.emz, p.d);
            }
            tile = J.tile_base + tz * p.tilesY + ty;
        }
        // One slot per candidate (no global counter); misses are skipped
        // by the scatter pass.
        if (c < total) pairs[c] = make_uint2(t, hit ? tile : UINT32_MAX);
        __syncthreads();
        uint32_t rank;
        if (hit) agg_insert(agg, tile, rank);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kAggSlots; i += blockDim.x)
            if (agg.key[i] != kAggEmpty) atomicAdd(&tile_hist[agg.key[i]], agg.cnt[i]);
    }
}
template __global__ void k_solid_candidates<false>(const Job *, const uint32_t *, uint32_t,
                                                   const float4 *, const uint32_t *,
                                                   const uint32_t *, const uint32_t *,
                                                   const uint4 *, uint2 *, uint16_t *,
                                                   uint32_t *, uint32_t, uint32_t *);
template __global__ void k_solid_candidates<true>(const Job *, const uint32_t *, uint32_t,
                                                  const float4 *, const uint32_t *,
                                                  const uint32_t *, const uint32_t *,
                                                  const uint4 *, uint2 *, uint16_t *,
                                                  uint32_t *, uint32_t, uint32_t *);

__global__ __launch_bounds__(256) void k_solid_scatter(const uint2 *__restrict__ pairs,
                                                       const uint16_t *__restrict__ pmask,
                                                       const uint32_t *__restrict__ cand_total,
                                                       const uint32_t *__restrict__ tile_off,
                                                       uint32_t *__restrict__ tile_cursor,
                                                       uint32_t *__restrict__ list,
                                                       uint16_t *__restrict__ list_mask,
                                                       uint32_t cap) {
    __shared__ AggTable agg;
    const uint32_t n = cand_total[1] ? 0u : min(*cand_total, cap);
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        __syncthreads();
        agg_clear(agg);
        __syncthreads();
        const uint32_t i = base + threadIdx.x;
        uint2 pr = make_uint2(0u, UINT32_MAX);
        uint32_t slot = 0, rank = 0;
        if (i < n) pr = pairs[i];
        const bool live = pr.y != UINT32_MAX;
        if (live) slot = agg_insert(agg, pr.y, rank);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kAggSlots; k += blockDim.x)
            if (agg.key[k] != kAggEmpty) agg.base[k] = atomicAdd(&tile_cursor[agg.key[k]], agg.cnt[k]);
        __syncthreads();
        if (live) {
            const uint32_t pos = tile_off[pr.y] + agg.base[slot] + rank;
            list[pos] = pr.x;
            if (list_mask) list_mask[pos] = pmask[i];
        }
    }
}

// One wave per 4x4 tile of voxel-centre rows.  Every queued triangle's
// crossings toggle one bit each in an LDS copy of the tile's 16 rows
// (ds_xor); a per-row prefix XOR then turns crossings into solid runs
// (the suffix flips of src/voxelizer.cu:1487-1616) and each bit word is
// written to HBM once, coalesced, with no global atomics.
__global__ __launch_bounds__(kWave) void k_solid_fill(const Job *__restrict__ jobs,
                                                      const uint32_t *__restrict__ tile_bases,
                                                      uint32_t njobs,
                                                      const float4 *__restrict__ rec,
                                                      const uint32_t *__restrict__ tile_off,
                                                      const uint32_t *__restrict__ tile_hist,
                                                      const uint32_t *__restrict__ list,
                                                      const uint16_t *__restrict__ list_mask) {
    extern __shared__ __align__(16) uint32_t cross[];  // [16][SW]
    const Job &J = jobs[job_of(tile_bases, njobs, blockIdx.x)];
    const SolidParams p = solid_params(J, rec);
    VK_G uint32_t *__restrict__ bits = gptr(J.bits);
    const uint32_t W = p.X >> 5;
    // Rows wider than kFillSegWords words are filled a segment at a time:
    // the tile's triangles are walked once per segment, and each row's
    // crossing parity carries from one segment into the next.
    const uint32_t SW = min(W, kFillSegWords);
    const uint32_t tid = threadIdx.x;
    const uint32_t local = blockIdx.x - J.tile_base;
    const uint32_t ty = local % p.tilesY;
    const uint32_t tz = local / p.tilesY;
    const uint32_t tile = blockIdx.x;
    const uint32_t cnt = tile_hist[tile];
    const uint32_t off = tile_off[tile];
    // The tile's 16 rows are split over 2^lg lanes per triangle (the largest
    // lg <= 4 with cnt << lg <= 64), so a tile with few triangles (the
    // median one on large meshes has 8) still keeps the wave's lanes busy;
    // crossings are XORed, so the split does not change the result.
    uint32_t lg = 0;
    while (lg < 4u && cnt <= (kWave >> (lg + 1u))) lg++;
    const uint32_t items = cnt << lg, rows = 16u >> lg;
    const uint32_t lane = tid;
    uint32_t carries = 0u;  // bit r: crossing parity of row r before the segment
    for (uint32_t s0 = 0; s0 < W; s0 += SW) {
        const uint32_t sw = min(SW, W - s0);
        uint4 *cross4 = reinterpret_cast<uint4 *>(cross);  // 16 * SW words: 4 * SW uint4
        for (uint32_t i = tid; i < 4u * SW; i += kWave) cross4[i] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
        for (uint32_t i = tid; i < items; i += kWave) {
            const uint32_t t = list[off + (i >> lg)];
            // Rows whose subspace queued the triangle (all rows with one subspace).
            const uint32_t rmask = list_mask ? (uint32_t)list_mask[off + (i >> lg)] : 0xFFFFu;
            const uint32_t r0 = (i & ((1u << lg) - 1u)) * rows;
            D3 v[3];
            load_tri_d(p, t, v);
            FillSetup fs;
            fill_setup(v, fs);
            for (uint32_t r = r0; r < r0 + rows; r++) {
                const uint32_t cy = ty * 4u + (r & 3u), cz = tz * 4u + (r >> 2);
                if (cy < p.smin[0] || cy >= p.smax[0] || cz < p.smin[1] || cz >= p.smax[1]) continue;
                if (!((rmask >> r) & 1u)) continue;
                uint32_t vx;
                if (fill_center(fs, p.emx, p.emy, p.emz, p.d, cy, cz, vx) && vx < p.X) {
                    const uint32_t wx = (vx >> 5) - s0;  // wraps for words before s0
                    if (wx < sw) atomicXor(&cross[r * SW + wx], 1u << (31u - (vx & 31u)));
                }
            }
        }
        __syncthreads();
        // Crossing parity -> solid bits (MSB-first bits: x%32 == 0 is bit
        // 31): a prefix XOR inside each word, and across words the parity of
        // the crossings before it (one ballot per 64 words), written
        // straight out.
        for (uint32_t r = 0; r < 16u; r++) {
            const uint32_t cy = ty * 4u + (r & 3u), cz = tz * 4u + (r >> 2);
            if (cy < p.smin[0] || cy >= p.smax[0] || cz < p.smin[1] || cz >= p.smax[1]) continue;
            const size_t rowIdx = (size_t)(cz + p.oz) * p.rowsY + (cy + p.oy);
            uint32_t carry = (carries >> r) & 1u;
            for (uint32_t w0 = 0; w0 < sw; w0 += kWave) {
                const uint32_t w = w0 + lane;
                const uint32_t c = w < sw ? cross[r * SW + w] : 0u;
                uint32_t sx = c;
                sx ^= sx >> 1;
                sx ^= sx >> 2;
                sx ^= sx >> 4;
                sx ^= sx >> 8;
                sx ^= sx >> 16;
                const unsigned long long odd = __ballot((__popc(c) & 1) != 0);
                // Odd-crossing words below this lane, counted by mbcnt.
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(odd >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)odd, 0u));
                const uint32_t before = carry ^ (below & 1u);
                if (w < sw) bits[rowIdx * W + s0 + w] = sx ^ (before ? 0xFFFFFFFFu : 0u);
                carry ^= (uint32_t)__popcll(odd) & 1u;
            }
            carries = (carries & ~(1u << r)) | (carry << r);
        }
        __syncthreads();  // the next segment clears cross[]
    }
}

__global__ __launch_bounds__(256) void k_zero_pad_rows(const Job *__restrict__ jobs,
                                                       const uint32_t *__restrict__ pad_bases,
                                                       uint32_t njobs) {
    const Job &J = jobs[job_of(pad_bases, njobs, blockIdx.x)];
    const uint32_t W = J.gd.W, rowsY = J.gd.Y, rowsZ = J.gd.Z;
    const uint32_t i = (blockIdx.x - J.pad_base) * 256u + threadIdx.x;
    if (i >= pad_words(W, rowsY, rowsZ)) return;
    const uint32_t r = i / W, w = i % W;
    uint32_t y, z, bit;
    if (r < 2u * rowsZ) {
        z = r >> 1;
        y = (r & 1u) ? rowsY - 1u : 0u;
        bit = (r & 1u) ? 2u : 1u;
    } else {
        const uint32_t q = r - 2u * rowsZ;
        y = q >> 1;
        z = (q & 1u) ? rowsZ - 1u : 0u;
        bit = (q & 1u) ? 8u : 4u;
    }
    // Rows and planes that are a neighbouring partition's overlap hold
    // voxelized data (written by the fill) instead of padding.
    if (J.padmask & bit) gptr(J.bits)[((size_t)z * rowsY + y) * W + w] = 0u;
}

// ---------------------------------------------------------------------------
// Surface stage
// ---------------------------------------------------------------------------

// Voxel range vb of a triangle clipped to the whole ext space (the union of
// all matSplitRes subspaces); false if it lies outside (adjustVoxelRange,
// src/voxelizer.cu:3963-3993).
__device__ __forceinline__ bool surf_clip(const SurfParams &p, VoxBB &vb) {
    if (0 > vb.max.x || 0 > vb.max.y || 0 > vb.max.z) return false;
    if (vb.min.x > p.res[0] - 1 || vb.min.y > p.res[1] - 1 || vb.min.z > p.res[2] - 1) return false;
    vb.min = {max(vb.min.x, 0), max(vb.min.y, 0), max(vb.min.z, 0)};
    vb.max = {min(vb.max.x, p.res[0] - 1), min(vb.max.y, p.res[1] - 1), min(vb.max.z, p.res[2] - 1)};
    return true;
}
__device__ __forceinline__ bool surf_range(const SurfParams &p, const F3 tri[3], VoxBB &vb) {
    F3 mn, mx;
    tri_bounds(tri, mn, mx);
    vb = voxel_bounds_half(mn, mx, p.m, p.d);
    return surf_clip(p, vb);
}

// Axis helpers: for dominant axis a, u and v are the other two axes in
// increasing order.
__device__ __forceinline__ int axis_u(int a) { return a == AX_X ? AX_Y : AX_X; }
__device__ __forceinline__ int axis_v(int a) { return a == AX_Z ? AX_Y : AX_Z; }
__device__ __forceinline__ float pick3(const F3 &p, int a) {
    return a == AX_X ? p.x : a == AX_Y ? p.y : p.z;
}
__device__ __forceinline__ int32_t pick3i(const I3 &p, int a) {
    return a == AX_X ? p.x : a == AX_Y ? p.y : p.z;
}

// Bricks are laid over alloc coordinates (ext + off).
__device__ __forceinline__ void brick_range(const SurfParams &p, const VoxBB &vb, uint32_t lo[3],
                                            uint32_t hi[3]) {
    lo[0] = (uint32_t)(vb.min.x + p.off[0]) / (uint32_t)p.bdim[0];
    lo[1] = (uint32_t)(vb.min.y + p.off[1]) / (uint32_t)p.bdim[1];
    lo[2] = (uint32_t)(vb.min.z + p.off[2]) / (uint32_t)p.bdim[2];
    hi[0] = (uint32_t)(vb.max.x + p.off[0]) / (uint32_t)p.bdim[0];
    hi[1] = (uint32_t)(vb.max.y + p.off[1]) / (uint32_t)p.bdim[1];
    hi[2] = (uint32_t)(vb.max.z + p.off[2]) / (uint32_t)p.bdim[2];
}

// Per triangle: class and dominant axis (classifyTriangles), and the box of
// bricks its clipped voxel range touches, packed as
// {info, lo.x | lo.y<<16, lo.z | w.x<<16, w.y | w.z<<16}.  With tri_ext,
// also its clipped voxel extents along (u, v, a) and its work items (voxels
// for 1D, dominant-plane columns otherwise), the cost model's inputs.
__global__ __launch_bounds__(256) void k_surf_setup(const Job *__restrict__ jobs,
                                                    const uint32_t *__restrict__ tri_bases,
                                                    uint32_t njobs, uint32_t nslots,
                                                    uint4 *__restrict__ tri_info,
                                                    uint32_t *__restrict__ cand_count,
                                                    uint4 *__restrict__ tri_ext,
                                                    float4 *__restrict__ tri_rec) {
    for (uint32_t t0 = blockIdx.x * 256u; t0 < nslots; t0 += gridDim.x * 256u) {
        const Job &J = jobs[job_of(tri_bases, njobs, t0)];
        const uint32_t t = t0 + threadIdx.x;
        if (t - J.tri_base >= J.nt || !J.use_mats) {  // padding, or no surface wanted
            cand_count[t] = 0u;
            continue;
        }
        const SurfParams p = surf_params(J);
        const float4 r0 = tri_rec[3 * (size_t)t], r1 = tri_rec[3 * (size_t)t + 1],
                     r2 = tri_rec[3 * (size_t)t + 2];
        const F3 tri[3] = {{r0.x, r0.y, r0.z}, {r0.w, r1.x, r1.y}, {r1.z, r1.w, r2.x}};
        // One voxel range for the class and for the brick box.
        F3 mn, mx;
        tri_bounds(tri, mn, mx);
        VoxBB vb = voxel_bounds_half(mn, mx, p.m, p.d);
        int bb, axis;
        classify_vb(tri, vb, bb, axis);
        uint32_t c = 0;
        uint32_t lo[3] = {0, 0, 0}, w[3] = {0, 0, 0};
        uint4 ext = make_uint4(0u, 0u, 0u, 0u);
        if (bb != BB_DEGEN && surf_clip(p, vb)) {
            uint32_t hi[3];
            brick_range(p, vb, lo, hi);
            for (int a = 0; a < 3; a++) w[a] = hi[a] - lo[a] + 1u;
            c = w[0] * w[1] * w[2];
            const uint32_t eu = (uint32_t)(pick3i(vb.max, axis_u(axis)) - pick3i(vb.min, axis_u(axis)) + 1);
            const uint32_t ev = (uint32_t)(pick3i(vb.max, axis_v(axis)) - pick3i(vb.min, axis_v(axis)) + 1);
            const uint32_t ea = (uint32_t)(pick3i(vb.max, axis) - pick3i(vb.min, axis) + 1);
            ext = make_uint4(eu, ev, ea, bb == BB_1D ? ea : eu * ev);
        }
        const uint32_t info = ((uint32_t)bb << 30) | ((uint32_t)axis << 28);
        tri_info[t] = make_uint4(info, lo[0] | (lo[1] << 16), lo[2] | (w[0] << 16),
                                 w[1] | (w[2] << 16));
        // The class word goes into the triangle's record for the brick pass.
        reinterpret_cast<float *>(tri_rec)[12 * (size_t)t + 9] = __uint_as_float(info);
        cand_count[t] = c;
        if (tri_ext) tri_ext[t] = ext;
    }
}

// One thread per (triangle, brick) candidate: counting pass (scatter = 0)
// or placement pass (scatter = 1) of the counting sort by brick.  The
// triangle of candidate c is found by binary search in the candidate offsets,
// narrowed per block to the triangles its 256 candidates span.
template <bool kMulti>
__global__ __launch_bounds__(256) void k_surf_bin(
    const Job *__restrict__ jobs, const uint32_t *__restrict__ tri_job, uint32_t nslots,
    const uint32_t *__restrict__ cand_off, const uint32_t *__restrict__ cand_total,
    const uint32_t *__restrict__ first, const uint4 *__restrict__ tri_info,
    uint32_t *__restrict__ brick_hist,
    const uint32_t *__restrict__ brick_off, uint32_t *__restrict__ brick_cursor,
    uint32_t *__restrict__ blist, int scatter, uint32_t cap, uint32_t *__restrict__ ovf) {
    __shared__ uint32_t trange[2];
    __shared__ AggTable agg;
    const uint32_t full = *cand_total, total = cand_total[1] ? 0u : min(full, cap);
    if (ovf && (full > cap || cand_total[1]) && blockIdx.x == 0 && threadIdx.x == 0) *ovf = 1u;
    for (uint32_t base = blockIdx.x * blockDim.x; base < total; base += gridDim.x * blockDim.x) {
        __syncthreads();
        agg_clear(agg);
        if (threadIdx.x < 2) {  // the triangles of candidates base and base + 256 bound the block's
            const uint32_t b = base / 256u + threadIdx.x;
            trange[threadIdx.x] = threadIdx.x == 0 || base + 256u < total ? first[b] : nslots - 1u;
        }
        __syncthreads();
        const uint32_t c = base + threadIdx.x;
        const bool live = c < total;
        uint32_t lo = trange[0], hi = trange[1] + 1u;
        if (!live) lo = hi = 0u;
        while (lo + 1u < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cand_off[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint32_t t = lo;
        uint32_t b = 0, slot = 0, rank = 0;
        if (live) {
            const Job &J = jobs[kMulti ? tri_job[t] : 0u];
            const uint4 ti = tri_info[t];
            const uint32_t bx0 = ti.y & 0xFFFFu, by0 = ti.y >> 16, bz0 = ti.z & 0xFFFFu;
            const uint32_t w0 = ti.z >> 16, w1 = ti.w & 0xFFFFu;
            const uint32_t local = c - cand_off[t];
            const uint32_t bx = bx0 + local % w0;
            const uint32_t by = by0 + (local / w0) % w1;
            const uint32_t bz = bz0 + local / (w0 * w1);
            b = J.brick_base + (bz * J.nb[1] + by) * J.nb[0] + bx;
            slot = agg_insert(agg, b, rank);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < kAggSlots; k += blockDim.x) {
            if (agg.key[k] == kAggEmpty) continue;
            if (!scatter)
                atomicAdd(&brick_hist[agg.key[k]], agg.cnt[k]);
            else
                agg.base[k] = atomicAdd(&brick_cursor[agg.key[k]], agg.cnt[k]);
        }
        __syncthreads();
        if (live && scatter) blist[brick_off[b] + agg.base[slot] + rank] = t;
    }
}
template __global__ void k_surf_bin<false>(const Job *, const uint32_t *, uint32_t,
                                           const uint32_t *, const uint32_t *, const uint32_t *,
                                           const uint4 *, uint32_t *, const uint32_t *,
                                           uint32_t *, uint32_t *, int, uint32_t, uint32_t *);
template __global__ void k_surf_bin<true>(const Job *, const uint32_t *, uint32_t,
                                          const uint32_t *, const uint32_t *, const uint32_t *,
                                          const uint4 *, uint32_t *, const uint32_t *,
                                          uint32_t *, uint32_t *, int, uint32_t, uint32_t *);

// Winner key of a (triangle, voxel) hit; the largest key of a voxel wins
// (bbType first, then the mesh-local triangle index, as the oracle's rule).
// With `embed` (meshes of fewer than 2^22 triangles) the key carries the
// triangle's material in its low byte (the index is unique, so the byte
// never decides), and the brick pass reads no materials from memory; larger
// meshes keep the plain index and look the material up.
__device__ __forceinline__ uint32_t winner_key(uint32_t bb, uint32_t lt, uint32_t mat,
                                               bool embed) {
    return (bb << 30) | (embed ? (lt << 8) | (mat & 0xFFu) : lt);
}

// Per-triangle data of one brick's chunk, set up once in LDS.  Coordinates
// are named (u, v, a): a is the triangle's dominant axis, u and v the other
// two in increasing axis order (u = x unless a = x; v = z unless a = z).
// Every 2-term edge function of the reference, (ne.x*c0 + ne.y*c1) + de, is
// kept as (c_first*p_first + c_second*p_second) + de with the two products
// in either order: IEEE addition is commutative, so the value is identical.
struct SurfTri {
    float dom[3][3];    // dominant-plane edges: (coef of u, coef of v, de)
    float pl[2][3][3];  // the two other planes (3D): (coef of a, coef of other, de)
    float n[3], t0[3];  // normal, first vertex (x, y, z)
    int32_t lo[3], hi[3];      // brick-clipped voxel range, (u, v, a) order
    int32_t vmin_a, vmax_a;    // unclipped range along a
    int32_t s[3];              // LDS key strides of (u, v, a)
    int32_t base;              // LDS key index of (lo_u, lo_v, lo_a)
    int32_t wu;                // columns along u
    int32_t ngu;               // work items along u (kSurfCols columns each)
    float rngu;                // 1 / ngu
    uint32_t key;              // LDS winner key (winner_key)
    uint32_t flags;            // bb | axis<<2 | (plane0 other is v)<<4 | (plane1 other is v)<<5
};
constexpr int kSurfChunk = 64;
// Columns (or 1D voxels) per work item.  Grouping neighbouring columns
// shares the item's lookup and v terms but lowers the parallelism the LDS
// latency needs: on the large hall 1 column is fastest (2 and 4 measured
// 9% and 13% slower).
constexpr int kSurfCols = 1;


// Sets up triangle t for the brick whose alloc-coordinate box starts at
// (ext coordinates) (blx, bly, blz), restricted to its z layers [z0, z1)
// (a sub-brick); returns its work items there: voxels along the axis for 1D,
// dominant-plane columns for 2D/3D.
__device__ __forceinline__ uint32_t surf_tri_setup(const SurfParams &p,
                                                   const float4 *__restrict__ tri_rec, uint32_t t,
                                                   uint32_t tri_base, bool embed,
                                                   int32_t blx, int32_t bly, int32_t blz,
                                                   int32_t z0, int32_t z1, SurfTri &st) {
    const float4 r0 = tri_rec[3 * (size_t)t], r1 = tri_rec[3 * (size_t)t + 1],
                 r2 = tri_rec[3 * (size_t)t + 2];
    const F3 tri[3] = {{r0.x, r0.y, r0.z}, {r0.w, r1.x, r1.y}, {r1.z, r1.w, r2.x}};
    const uint32_t info = __float_as_uint(r2.y);
    VoxBB vb;
    if (!surf_range(p, tri, vb)) return 0u;
    const int bb = (int)(info >> 30), a = (int)((info >> 28) & 3u);
    const int ua = axis_u(a), va = axis_v(a);
    const int32_t bd[3] = {p.bdim[0], p.bdim[1], p.bdim[2]};
    const int32_t stride[3] = {1, bd[0], bd[0] * bd[1]};
    const int ax3[3] = {ua, va, a};
    int32_t lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int ax = ax3[k];
        const int32_t bl = ax == AX_X ? blx : ax == AX_Y ? bly : blz + z0;
        const int32_t bs = ax == AX_X ? bd[0] : ax == AX_Y ? bd[1] : z1 - z0;
        lo[k] = max(pick3i(vb.min, ax), bl);
        hi[k] = min(pick3i(vb.max, ax), bl + bs - 1);
        if (lo[k] > hi[k]) return 0u;
        st.lo[k] = lo[k];
        st.hi[k] = hi[k];
        st.s[k] = ax == AX_X ? stride[0] : ax == AX_Y ? stride[1] : stride[2];
    }
    st.base = (lo[0] - (ua == AX_X ? blx : ua == AX_Y ? bly : blz)) * st.s[0] +
              (lo[1] - (va == AX_X ? blx : va == AX_Y ? bly : blz)) * st.s[1] +
              (lo[2] - (a == AX_X ? blx : a == AX_Y ? bly : blz)) * st.s[2];
    st.vmin_a = pick3i(vb.min, a);
    st.vmax_a = pick3i(vb.max, a);
    st.key = winner_key((uint32_t)bb, t - tri_base, __float_as_uint(r2.z), embed);
    st.wu = hi[0] - lo[0] + 1;
    st.ngu = (st.wu + kSurfCols - 1) / kSurfCols;
    st.rngu = 1.0f / (float)st.ngu;
    uint32_t flags = (uint32_t)bb | ((uint32_t)a << 2);
    if (bb == BB_1D) {
        st.flags = flags;
        return (uint32_t)((hi[2] - lo[2] + kSurfCols) / kSurfCols);
    }
    const F3 n = fcross(fsub(tri[0], tri[2]), fsub(tri[1], tri[0]));
    st.n[0] = n.x; st.n[1] = n.y; st.n[2] = n.z;
    st.t0[0] = tri[0].x; st.t0[1] = tri[0].y; st.t0[2] = tri[0].z;
    // Plane P tests coordinates (c0, c1): plane 0 (y, z), 1 (z, x), 2 (x, y).
    auto first_axis = [](int P) { return P == 0 ? AX_Y : P == 1 ? AX_Z : AX_X; };
    auto second_axis = [](int P) { return P == 0 ? AX_Z : P == 1 ? AX_X : AX_Y; };
    {
        OverlapData o;
        setup_overlap(tri, n, p.d, a, o);
        const bool first_is_u = first_axis(a) == ua;
#pragma unroll
        for (int e = 0; e < 3; e++) {
            st.dom[e][0] = first_is_u ? o.ne[e].x : o.ne[e].y;
            st.dom[e][1] = first_is_u ? o.ne[e].y : o.ne[e].x;
            st.dom[e][2] = o.de[e];
        }
    }
    if (bb == BB_3D) {
        const int planes[2] = {a == AX_X ? 1 : 0, a == AX_Z ? 1 : 2};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            OverlapData o;
            setup_overlap(tri, n, p.d, planes[q], o);
            const bool first_is_a = first_axis(planes[q]) == a;
            const int other = first_is_a ? second_axis(planes[q]) : first_axis(planes[q]);
            if (other == va) flags |= 1u << (4 + q);
#pragma unroll
            for (int e = 0; e < 3; e++) {
                st.pl[q][e][0] = first_is_a ? o.ne[e].x : o.ne[e].y;
                st.pl[q][e][1] = first_is_a ? o.ne[e].y : o.ne[e].x;
                st.pl[q][e][2] = o.de[e];
            }
        }
    }
    st.flags = flags;
    return (uint32_t)(st.ngu * (hi[1] - lo[1] + 1));
}

__device__ __forceinline__ void key_write(uint32_t *keys, uint8_t *touched, int32_t idx,
                                          int32_t nvox, int32_t row_shift, uint32_t key) {
    if ((uint32_t)idx >= (uint32_t)nvox) return;  // unreachable: items lie in the brick
    atomicMax(&keys[idx], key);
    touched[idx >> row_shift] = 1;
}

// One work item, following the 1D/2D/3D kernels of the reference
// (src/voxelizer.cu:3307-3955) with their matSplitRes subspace clipping --
// only the 3D single-voxel shortcut depends on the subspace, through the
// dominant-axis interval the voxel falls in.
__device__ __forceinline__ void surf_item(const SurfParams &p, const SplitTable &spl, const SurfTri &st,
                          uint32_t k, uint32_t *keys, uint8_t *touched, int32_t nvox,
                          int32_t row_shift) {
    const uint32_t flags = st.flags;
    const int bb = (int)(flags & 3u), a = (int)((flags >> 2) & 3u);
    if (bb == BB_1D) {
        const int32_t len = st.hi[2] - st.lo[2] + 1;
        for (int j = 0; j < kSurfCols; j++) {
            const int32_t kk = (int32_t)k * kSurfCols + j;
            if (kk >= len) break;
            key_write(keys, touched, st.base + (int32_t)__umul24(kk, st.s[2]), nvox, row_shift,
                      st.key);
        }
        return;
    }
    // item k covers columns u0 .. u0+kSurfCols-1 of row v:
    // k = (v - lo_v) * ngu + (u0 - lo_u) / kSurfCols
    // Strides and indices here are below 2^24 (bricks hold at most 4096
    // voxels), so products use the full-rate 24-bit multiply.
    int32_t q = (int32_t)((float)k * st.rngu);
    int32_t r = (int32_t)k - (int32_t)__umul24(q, st.ngu);
    if (r < 0) { q--; r += st.ngu; }
    if (r >= st.ngu) { q++; r -= st.ngu; }
    const int32_t v = st.lo[1] + q;
    const int ua = axis_u(a), va = axis_v(a);
    const float mx = p.m.x, my = p.m.y, mz = p.m.z;
    const float mu = ua == AX_X ? mx : (ua == AX_Y ? my : mz);
    const float mv = va == AX_X ? mx : (va == AX_Y ? my : mz);
    const float ma = a == AX_X ? mx : (a == AX_Y ? my : mz);
    const float pv = vox_corner(v, mv, p.d);
    float kdom[3];  // v terms of the dominant-plane edge functions
#pragma unroll
    for (int e = 0; e < 3; e++) kdom[e] = st.dom[e][1] * pv;
    const bool threeD = bb == BB_3D;
    const F3 t0 = {st.t0[0], st.t0[1], st.t0[2]}, nn = {st.n[0], st.n[1], st.n[2]};
    const uint32_t ns = spl.n[a];
    for (int j = 0; j < kSurfCols; j++) {
        const int32_t cu = r * kSurfCols + j;  // column offset along u
        if (cu >= st.wu) break;
        const int32_t u = st.lo[0] + cu;
        const float pu = vox_corner(u, mu, p.d);
        bool in = true;
#pragma unroll
        for (int e = 0; e < 3; e++) in &= (st.dom[e][0] * pu + kdom[e]) + st.dom[e][2] >= 0.0f;
        if (!in) continue;
        const int32_t col =
            st.base + (int32_t)__umul24(cu, st.s[0]) + (int32_t)__umul24(q, st.s[1]);
        if (!threeD) {
            key_write(keys, touched, col, nvox, row_shift, st.key);
            continue;
        }
        const F3 pp = {a == AX_X ? 0.0f : pu, a == AX_Y ? 0.0f : (a == AX_X ? pu : pv),
                       a == AX_Z ? 0.0f : pv};
        int32_t lo, hi;
        depth_range(t0, pp, nn, p.m, p.d, a, lo, hi);
        const int32_t ilo = max(lo, st.lo[2]), ihi = min(hi, st.hi[2]);
        if (ilo > ihi) continue;
        // The two side planes' edge functions along the column,
        // (c * pa + k) + e, with their constant parts in registers.
        float sc[6], sk[6], se[6];
#pragma unroll
        for (int q2 = 0; q2 < 2; q2++) {
            const float po = (flags >> (4 + q2)) & 1u ? pv : pu;
#pragma unroll
            for (int e = 0; e < 3; e++) {
                sc[3 * q2 + e] = st.pl[q2][e][0];
                sk[3 * q2 + e] = st.pl[q2][e][1] * po;
                se[3 * q2 + e] = st.pl[q2][e][2];
            }
        }
        // matSplitRes interval of the voxel along a (advanced as i grows)
        // and the single-voxel shortcut's range clipped to it.
        uint32_t si = 0;
        while (si + 1 < ns && (uint32_t)ilo >= spl.hi[a][si]) si++;
        uint32_t shi = spl.hi[a][si];
        int32_t rl = max(lo, max(st.vmin_a, (int32_t)spl.lo[a][si]));
        int32_t rh = min(hi, min(st.vmax_a, (int32_t)shi - 1));
        const int32_t sa = st.s[2], col0 = col - __mul24(st.lo[2], sa);  // lo may be -1
        for (int32_t i = ilo; i <= ihi; i++) {
            if (si + 1 < ns && (uint32_t)i >= shi) {
                do si++;
                while (si + 1 < ns && (uint32_t)i >= spl.hi[a][si]);
                shi = spl.hi[a][si];
                rl = max(lo, max(st.vmin_a, (int32_t)spl.lo[a][si]));
                rh = min(hi, min(st.vmax_a, (int32_t)shi - 1));
            }
            bool write = rl == rh;
            if (!write) {
                const float pa = vox_corner(i, ma, p.d);
                float f[6];
#pragma unroll
                for (int e = 0; e < 6; e++) f[e] = (sc[e] * pa + sk[e]) + se[e];
                write = f[0] >= 0.0f && f[1] >= 0.0f && f[2] >= 0.0f && f[3] >= 0.0f &&
                        f[4] >= 0.0f && f[5] >= 0.0f;
            }
            if (write) key_write(keys, touched, col0 + __mul24(i, sa), nvox, row_shift, st.key);
        }
    }
}

// Persistent workgroups over a list of non-empty bricks (`order`: brick ids
// in index order for the plain schedule, or sorted by predicted cost).  Each
// workgroup takes runs of list entries from a global counter (the grab for
// the following run is issued while the last brick of the current one is
// processed), so bricks of uneven cost balance over the GPU.  Per brick: its triangles are
// taken 64 at a time, each set up once into LDS, and all (triangle, column)
// items of the chunk are spread over the 256 threads; winners accumulate in
// LDS (ds_max of bbType<<30 | triangle).  Only rows of the brick that
// received a key are visited afterwards: one ballot per 64 voxels gives the
// surface bit words, one lane per surface voxel fetches its winner's material
// into the brick's block of the material grid (one byte per voxel, brick
// order), and the rows' keys are cleared for the next brick.  Surface bits
// are ORed into the solid grid.  With `stats`, surface voxels are counted;
// with `cycles`, each brick's duration (s_memtime) is recorded.  The kKeys
// variant also stores the winner keys of jobs with a keygrid (PartialNode).
// DPP controls of the epilogue's 8-lane OR (all lanes active).
constexpr int kDppQuadSwap1 = 0xB1, kDppQuadSwap2 = 0x4E, kDppHalfMirror = 0x141;

template <bool kKeys>
__global__ __launch_bounds__(256, 5) void k_surf_brick(
    const Job *__restrict__ jobs, const uint32_t *__restrict__ brick_bases, uint32_t njobs,
    const float4 *__restrict__ tri_rec, const uint32_t *__restrict__ brick_off,
    const uint32_t *__restrict__ brick_hist, const uint32_t *__restrict__ blist,
    uint32_t *__restrict__ stats, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ n_order, const uint32_t *__restrict__ hlist,
    const uint32_t *__restrict__ hsub, const uint32_t *__restrict__ n_heavy,
    const uint8_t *__restrict__ skip, uint32_t *__restrict__ next,
    uint64_t *__restrict__ cycles) {
    __shared__ __align__(16) uint32_t keys[kMaxBrickVox];
    __shared__ SurfTri tris[kSurfChunk];
    __shared__ uint32_t cstart[kSurfChunk], ctotal[2], smask[256 / kWave][2];
    __shared__ uint8_t cslot[kSurfChunk];
    __shared__ uint32_t wsurf[kMaxBrickVox / 32];
    __shared__ uint8_t touched[32];
    __shared__ uint32_t s_q, s_qe;
    __shared__ SplitTable spl;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
    for (int32_t i = tid; i < kMaxBrickVox; i += blockDim.x) keys[i] = 0u;
    if (tid < 32u) touched[tid] = 0;
    if (tid < 2u * (256u / kWave)) (&smask[0][0])[tid] = 0u;
    // Entries: the heavy list first (learned order), then the plain list.
    const uint32_t nh = hlist ? *n_heavy : 0u;
    const uint32_t n_list = *n_order + nh;
    // Guided grabs: a workgroup takes a run of list entries sized to the
    // work left (remaining / (8 * workgroups), 1..32; heavy entries one at a
    // time), so the shared counter sees a few grabs per workgroup rather
    // than one per brick (one address bumped per brick serialises the pass),
    // while the tail still goes one brick at a time.
    auto grab = [&](uint32_t seen, uint32_t &qe) {
        const uint32_t left = n_list > seen ? n_list - seen : 0u;
        const uint32_t k = seen < nh ? 1u : min(max(left / (8u * gridDim.x), 1u), 32u);
        const uint32_t q0 = atomicAdd(next, k);
        qe = min(q0 + k, n_list);
        return q0;
    };
    if (tid == 0) {
        uint32_t qe;
        s_q = grab(0u, qe);
        s_qe = qe;
    }
    __syncthreads();
    uint32_t q = s_q, qend = s_qe;
    uint32_t cur_job = UINT32_MAX;  // job whose split table is in LDS
    uint32_t nsurf = 0u;            // surface voxels found by this lane
    while (q < n_list) {  // q and qend are uniform over the block
        uint32_t q_next = q + 1u, qe_next = qend;
        if (tid == 0 && q + 1u >= qend) q_next = grab(qend, qe_next);  // next run, early
        // The brick's header comes through uniform (scalar) loads, which do
        // not wait behind the previous brick's outstanding vector stores.
        const uint32_t qs = __builtin_amdgcn_readfirstlane(q);
        const bool from_heavy = qs < nh;
        const uint32_t gb =
            __builtin_amdgcn_readfirstlane(from_heavy ? hlist[qs] : order[qs - nh]);
        // A plain entry already taken from the heavy list is passed over
        // (its flag byte read as part of an aligned dword, uniformly).
        const bool skipped =
            !from_heavy && hlist &&
            ((__builtin_amdgcn_readfirstlane(
                  ((const uint32_t *)skip)[(qs - nh) >> 2]) >> (8u * ((qs - nh) & 3u))) & 0xFFu);
        // Layer range of the entry: a heavy entry may be a sub-brick.
        const uint32_t sub =
            from_heavy ? __builtin_amdgcn_readfirstlane(hsub[qs]) : 0xFFFF0000u;
        if (!skipped) {
            const uint32_t cnt = __builtin_amdgcn_readfirstlane(brick_hist[gb]);
            uint64_t t_start = 0;
            if (cycles && tid == 0) t_start = __builtin_amdgcn_s_memtime();
            // Consecutive entries mostly share a job: its range is tested
            // before a search.
            const uint32_t j =
                njobs > 1u && cur_job != UINT32_MAX && gb >= brick_bases[cur_job] &&
                        gb < brick_bases[cur_job + 1u]
                    ? cur_job
                    : job_of(brick_bases, njobs, gb);
            const Job &J = jobs[j];
            if (j != cur_job) {  // uniform: the job's matSplitRes table into LDS
                __syncthreads();
                const VK_G uint32_t *src = (const VK_G uint32_t *)gptr(J.splits);
                uint32_t *dst = (uint32_t *)&spl;
                for (uint32_t i = tid; i < sizeof(SplitTable) / 4; i += blockDim.x) dst[i] = src[i];
                cur_job = j;
            }
            const SurfParams p = surf_params(J);
            const VK_G uint8_t *__restrict__ materials = gptr(J.mats);
            const bool embed = J.embed != 0;
            const uint32_t b = gb - J.brick_base;  // brick within the job
            const int32_t bd0 = p.bdim[0], bd1 = p.bdim[1], bd2 = p.bdim[2];
            const int32_t nvox = bd0 * bd1 * bd2;
            const int32_t z0 = (int32_t)(sub & 0xFFFFu), z1 = min((int32_t)(sub >> 16), bd2);
            const int32_t wpr = bd0 / 32;     // words per row
            const int32_t row_shift = 31 - __clz(bd0);
            const uint32_t bxi = b % p.nb[0], byi = (b / p.nb[0]) % p.nb[1],
                           bzi = b / (p.nb[0] * p.nb[1]);
            // Bricks are laid over alloc coordinates; ext = alloc - off.
            const int32_t blx = (int32_t)(bxi * bd0) - p.off[0],
                          bly = (int32_t)(byi * bd1) - p.off[1],
                          blz = (int32_t)(bzi * bd2) - p.off[2];
            const int32_t nwords = nvox / 32;
            // This thread's brick word (for the OR into the solid grid), early;
            // a sub-brick owns the words of its layers only.
            size_t gi = 0;
            bool gword = false;
            uint32_t bw = 0u;
            if ((int32_t)tid < nwords) {
                const int32_t wx = (int32_t)tid % wpr, ly = ((int32_t)tid / wpr) % bd1,
                              lz = (int32_t)tid / (wpr * bd1);
                const uint32_t aw = bxi * (uint32_t)wpr + (uint32_t)wx;
                const uint32_t ay = byi * (uint32_t)bd1 + (uint32_t)ly;
                const uint32_t az = bzi * (uint32_t)bd2 + (uint32_t)lz;
                gword = lz >= z0 && lz < z1 && aw < p.alloc[0] / 32u && ay < p.alloc[1] &&
                        az < p.alloc[2];
                if (gword) {
                    gi = ((size_t)az * p.alloc[1] + ay) * (p.alloc[0] / 32u) + aw;
                    bw = gptr(J.bits)[gi];
                }
            }
            const uint32_t off = __builtin_amdgcn_readfirstlane(brick_off[gb]);
            for (uint32_t c0 = 0; c0 < cnt; c0 += kSurfChunk) {
                __syncthreads();
                if (tid < (uint32_t)kSurfChunk) {
                    uint32_t items = 0u;
                    if (c0 + tid < cnt) {
                        const uint32_t t = blist[off + c0 + tid];
                        items = surf_tri_setup(p, tri_rec, t, J.tri_base, embed, blx, bly, blz,
                                               z0, z1, tris[tid]);
                    }
                    static_assert(kSurfChunk == kWave, "one wave sets up a chunk");
                    const uint32_t incl = wave_incl_scan(items);
                    // The triangles with items, in order: first item and LDS slot.
                    const unsigned long long nz = __ballot(items != 0u);
                    if (items) {
                        const uint32_t r = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(nz >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nz, 0u));
                        cstart[r] = incl - items;
                        cslot[r] = (uint8_t)tid;
                    }
                    if (tid == (uint32_t)kSurfChunk - 1u) {
                        ctotal[0] = incl;
                        ctotal[1] = (uint32_t)__popcll(nz);
                    }
                }
                __syncthreads();
                const uint32_t total = ctotal[0], ntri = ctotal[1];
                // Lane j holds the first item of the j-th triangle with items
                // (strictly increasing in j) and its slot.  A wave takes 64
                // consecutive items at a time: the triangle of its first item
                // is a ballot over the held starts, and each lane adds the
                // triangles starting between the first item and its own, read
                // from a 64-bit mask of start offsets that the starting lanes
                // OR together in LDS.
                const uint32_t cs = lane < ntri ? cstart[lane] : UINT32_MAX;
                const uint32_t sl = lane < ntri ? (uint32_t)cslot[lane] : 0u;
                uint32_t *sm = &smask[wave][0];
                for (uint32_t w0 = wave * kWave; w0 < total; w0 += blockDim.x) {  // uniform
                    const uint32_t t0 = (uint32_t)__popcll(__ballot(cs <= w0)) - 1u;
                    const uint32_t ps = cs - w0 - 1u;  // start offset - 1, in 0..62
                    if (cs > w0 && ps < kWave - 1u) atomicOr(&sm[ps >> 5], 1u << (ps & 31u));
                    const uint32_t m_lo = sm[0], m_hi = sm[1];
                    if (lane == 0u) {
                        sm[0] = 0u;
                        sm[1] = 0u;
                    }
                    const uint32_t ti =
                        t0 + __builtin_amdgcn_mbcnt_hi(m_hi, __builtin_amdgcn_mbcnt_lo(m_lo, 0u));
                    const uint32_t slot = (uint32_t)__shfl((int)sl, (int)ti);
                    const uint32_t first = (uint32_t)__shfl((int)cs, (int)ti);
                    const uint32_t c = w0 + lane;
                    if (c < total)
                        surf_item(p, spl, tris[slot], c - first, keys, touched, nvox, row_shift);
                }
            }
            __syncthreads();
            // Surface bit words and winner materials, 256 voxels per wave
            // step: each lane takes 4 consecutive voxels (one 16-byte LDS
            // read of their keys), the 8 lanes of a 32-voxel word OR their
            // nibbles together, and each lane of a word holding a surface
            // voxel writes its 4 material bytes as one dword of the brick's
            // block in the material grid (the winner's material, 0 for the
            // word's other voxels).  Steps whose rows received no key only
            // zero their surface words.  Keys are cleared for the next brick.
            // With materials in the keys the loop issues no global loads, so
            // its stores never wait (a wait for a load also waits for earlier
            // stores here).
            VK_G uint32_t *__restrict__ mg =
                (VK_G uint32_t *)(gptr(J.matgrid) + (size_t)b * (uint32_t)nvox);
            VK_G u32x4 *__restrict__ kg =
                kKeys && J.keygrid ? (VK_G u32x4 *)(gptr(J.keygrid) + (size_t)b * (uint32_t)nvox)
                                   : nullptr;
            const uint32_t tmask = (uint32_t)__ballot(lane < 32u && touched[lane & 31u] != 0);
            const uint32_t rows_per_step = 256u >> row_shift;  // bd0 >= 64: 1, 2 or 4 rows
            auto rows_out = [&](auto embed_c) {
                constexpr bool kEmbedded = decltype(embed_c)::value;
                for (int32_t v0 = (int32_t)wave * 256; v0 < nvox; v0 += 1024) {
                    const uint32_t rmask = ((1u << rows_per_step) - 1u) << (v0 >> row_shift);
                    if (!(tmask & rmask)) {
                        if (lane < 8u) wsurf[(v0 >> 5) + (int32_t)lane] = 0u;
                        continue;
                    }
                    uint4 *kp = (uint4 *)&keys[v0 + 4 * (int32_t)lane];
                    const uint4 kv = *kp;
                    const uint32_t nib = (kv.x ? 8u : 0u) | (kv.y ? 4u : 0u) | (kv.z ? 2u : 0u) |
                                         (kv.w ? 1u : 0u);
                    // OR over the word's 8 lanes in registers (DPP): the
                    // quad swaps [1,0,3,2] and [2,3,0,1], then the half-row
                    // mirror (lane i <-> 7 - i) joins the two quads.
                    uint32_t w = nib << (28u - 4u * (lane & 7u));
                    w |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w, kDppQuadSwap1, 0xF, 0xF, false);
                    w |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w, kDppQuadSwap2, 0xF, 0xF, false);
                    w |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w, kDppHalfMirror, 0xF, 0xF, false);
                    if ((lane & 7u) == 0u) wsurf[(v0 >> 5) + (int32_t)(lane >> 3)] = w;
                    if (!w) continue;
                    nsurf += (uint32_t)__popc(nib);
                    uint32_t m;
                    if (kEmbedded) {
                        m = (kv.x & 0xFFu) | (kv.y & 0xFFu) << 8 | (kv.z & 0xFFu) << 16 |
                            (kv.w & 0xFFu) << 24;
                    } else {
                        auto mat = [&](uint32_t key) -> uint32_t {
                            return key && materials ? (uint32_t)materials[key & 0x3FFFFFFFu] : 0u;
                        };
                        m = mat(kv.x) | mat(kv.y) << 8 | mat(kv.z) << 16 | mat(kv.w) << 24;
                    }
                    if (nib) *kp = make_uint4(0u, 0u, 0u, 0u);
                    mg[(v0 >> 2) + (int32_t)lane] = m;
                    if (kKeys && kg) kg[(v0 >> 2) + (int32_t)lane] = u32x4{kv.x, kv.y, kv.z, kv.w};
                }
            };
            if (embed)
                rows_out(std::integral_constant<bool, true>{});
            else
                rows_out(std::integral_constant<bool, false>{});
            __syncthreads();
            const uint32_t sw = gword ? wsurf[tid] : 0u;
            if (gword) {
                if (J.surf) gptr(J.surf)[gi] = sw;
                if (sw) gptr(J.bits)[gi] = bw | sw;
            }
            // The brick's word mask (words of other layers belong to other
            // parts of a cut brick, hence the OR into a cleared mask).
            const unsigned long long hm = __ballot(sw != 0u);
            if (wave < 2u && lane == 0u && hm) {
                uint32_t *wm = J.wmask + (size_t)b * 4u + 2u * wave;
                if ((uint32_t)hm) atomicOr(wm, (uint32_t)hm);
                if ((uint32_t)(hm >> 32)) atomicOr(wm + 1, (uint32_t)(hm >> 32));
            }
            if (cycles && tid == 0) cycles[gb] = __builtin_amdgcn_s_memtime() - t_start;
        }  // brick
        __syncthreads();
        if (tid < 32u) touched[tid] = 0;
        if (tid == 0) {
            s_q = q_next;
            s_qe = qe_next;
        }
        __syncthreads();
        q = s_q;
        qend = s_qe;
    }
    // One atomic per wave at exit: a shared counter bumped per brick would
    // serialise every workgroup on one address.
    if (stats) {
        for (int d = kWave / 2; d >= 1; d >>= 1) nsurf += (uint32_t)__shfl_xor((int)nsurf, d);
        if (lane == 0 && nsurf) atomicAdd(stats, nsurf);
    }
}
template __global__ void k_surf_brick<false>(const Job *, const uint32_t *, uint32_t,
                                             const float4 *, const uint32_t *, const uint32_t *,
                                             const uint32_t *, uint32_t *, const uint32_t *,
                                             const uint32_t *, const uint32_t *, const uint32_t *,
                                             const uint32_t *, const uint8_t *, uint32_t *,
                                             uint64_t *);
template __global__ void k_surf_brick<true>(const Job *, const uint32_t *, uint32_t,
                                            const float4 *, const uint32_t *, const uint32_t *,
                                            const uint32_t *, uint32_t *, const uint32_t *,
                                            const uint32_t *, const uint32_t *, const uint32_t *,
                                            const uint32_t *, const uint8_t *, uint32_t *,
                                            uint64_t *);

// Non-empty bricks in index order: flags, then (after an exclusive scan of
// the flags) the compacted list.
__global__ __launch_bounds__(256) void k_nonzero_flags(const uint32_t *__restrict__ hist,
                                                       uint32_t *__restrict__ flags, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        flags[i] = hist[i] != 0u ? 1u : 0u;
}
__global__ __launch_bounds__(256) void k_nonzero_list(const uint32_t *__restrict__ hist,
                                                      const uint32_t *__restrict__ off,
                                                      uint32_t *__restrict__ list, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (hist[i] != 0u) list[off[i]] = i;
}

// ---------------------------------------------------------------------------
// Erosion: removal of solid voxels that lack a solid neighbour on some axis
// (the invalid-permutation removal of src/voxelizer.cu:1791-1913), repeated
// until a pass removes nothing (src/host_side.cpp:1183-1206).  The rule is
// monotone and only removes, so passes may run in place and in any order:
// every such (chaotic) iteration reaches the same greatest fixed point as
// the reference's.  The first pass of a subspace reads the whole grid; a
// later pass only visits blocks that changed, or that hold neighbours of
// words in blocks that changed, in the previous pass (stamp = 1 + last pass with a
// change), and a word is written only when it changes.  Voxels outside
// [y0,y1) x [z0,z1) are left alone.  Plane-linear launch: a block takes
// kChunk consecutive words of a plane over kErodeZ planes.
// ---------------------------------------------------------------------------
// Whether logical block GL of job J lies in the subspace and, after the
// subspace's first pass, is near a change of the previous pass: a word
// depends on words at most W positions away in its plane and on the same
// word in the planes next to it, so the block is visited when any chunk
// within W words, or this chunk in the plane groups next to it, changed.
// DPP controls: every lane takes the value of the lane one below (wave_shr:1)
// or one above (wave_shl:1); the wave's end lanes get 0.
constexpr int kDppWaveShl1 = 0x130, kDppWaveShr1 = 0x138;

__device__ __forceinline__ bool erode_visit(const Job &J, uint32_t GL, uint32_t y0, uint32_t y1,
                                            uint32_t z0, uint32_t z1,
                                            const uint32_t *__restrict__ stamp, uint32_t gpass,
                                            int full) {
    const GridDims g = J.gd;
    const uint32_t P = g.W * g.Y;  // words per plane
    const uint32_t nch = (P + kChunk - 1u) / kChunk;
    const uint32_t nzc = (g.Z + kErodeZ - 1) / kErodeZ;
    const uint32_t L = GL - J.erode_base;  // block within the job
    stamp += J.erode_base;
    const uint32_t c = L % nch, zg = L / nch;
    const uint32_t zb = zg * kErodeZ;
    const uint32_t i0 = c * kChunk;
    const uint32_t ylo = i0 / g.W, yhi = min(i0 + kChunk - 1u, P - 1u) / g.W;
    if (zb >= z1 || zb + kErodeZ <= z0 || yhi < y0 || ylo >= y1) return false;
    if (full) return true;
    const uint32_t r = g.W / kChunk + 1u;
    const uint32_t clo = c >= r ? c - r : 0u, chi = min(c + r, nch - 1u);
    bool visit = false;
    for (uint32_t cc = clo; cc <= chi; cc++) visit |= stamp[zg * nch + cc] >= gpass;
    if (zg > 0) visit |= stamp[L - nch] >= gpass;
    if (zg + 1 < nzc) visit |= stamp[L + nch] >= gpass;
    return visit;
}

// One logical block GL (of job j) of an erosion pass.
__device__ __forceinline__ void erode_block(const Job &J, uint32_t j, uint32_t GL, uint32_t njobs,
                                            uint32_t y0, uint32_t y1, uint32_t z0, uint32_t z1,
                                            uint32_t *__restrict__ stamp,
                                            uint32_t *__restrict__ changes,
                                            uint32_t *__restrict__ totals,
                                            uint32_t *__restrict__ notconv, uint32_t pass,
                                            uint32_t gpass) {
    const GridDims g = J.gd;
    VK_G uint32_t *__restrict__ bits = gptr(J.bits);
    const uint32_t P = g.W * g.Y;  // words per plane
    const uint32_t nch = (P + kChunk - 1u) / kChunk;
    const uint32_t L = GL - J.erode_base;  // block within the job
    stamp += J.erode_base;
    const uint32_t c = L % nch, zg = L / nch;
    const uint32_t zb = zg * kErodeZ;
    const uint32_t i0 = c * kChunk;
    const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1u);
    const uint32_t i = i0 + tid;
    const bool valid = i < P;
    const uint32_t y = valid ? i / g.W : 0u;
    const uint32_t w = valid ? i - y * g.W : 0u;
    const size_t rowW = g.W, planeW = P;
    const size_t col0 = valid ? i : 0;  // index of (w, y) in plane 0
    const bool hyn = valid && y > 0, hyp = valid && y + 1 < g.Y;
    // Column z = zb-1 .. zb+kErodeZ of this word, the rows above and below it
    // for the thread's planes, and the words across the wave's edges: all
    // loads are independent and in flight together (clamped addresses).
    uint32_t col[kErodeZ + 2], yn[kErodeZ], yp[kErodeZ], em[kErodeZ], ep[kErodeZ];
#pragma unroll
    for (uint32_t k = 0; k < kErodeZ + 2; k++) {
        const uint32_t z = zb + k - 1u;  // wraps to UINT32_MAX for zb == 0, k == 0
        const bool ok = valid && z < g.Z;
        col[k] = bits[ok ? (size_t)z * planeW + col0 : 0];
        if (!ok) col[k] = 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kErodeZ; k++) {
        const uint32_t z = zb + k;
        const bool ok = valid && z < g.Z;
        const size_t wi = ok ? (size_t)z * planeW + col0 : 0;
        yn[k] = bits[ok && hyn ? wi - rowW : 0];
        yp[k] = bits[ok && hyp ? wi + rowW : 0];
        if (!(ok && hyn)) yn[k] = 0u;
        if (!(ok && hyp)) yp[k] = 0u;
        const bool lm = ok && lane == 0 && w > 0;
        const bool rp = ok && lane == kWave - 1 && w + 1 < g.W;
        em[k] = lm ? bits[wi - 1] : 0u;
        ep[k] = rp ? bits[wi + 1] : 0u;
    }
    bool changed = false;
#pragma unroll
    for (uint32_t k = 0; k < kErodeZ; k++) {
        const uint32_t z = zb + k;
        const uint32_t G = col[k + 1];
        // x neighbours: adjacent lanes hold the adjacent words of the plane
        // (whole-wave DPP shifts, no LDS round trip); across a row's ends
        // there is nothing.
        uint32_t gm = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)G, kDppWaveShr1, 0xF, 0xF, false);
        uint32_t gp = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)G, kDppWaveShl1, 0xF, 0xF, false);
        if (lane == 0) gm = em[k];
        if (lane == kWave - 1) gp = ep[k];
        if (w == 0) gm = 0u;
        if (w + 1 >= g.W) gp = 0u;
        if (!valid || !G || z >= g.Z || z < z0 || z >= z1 || y < y0 || y >= y1) continue;
        const uint32_t Lm = (G >> 1) | (gm << 31);
        const uint32_t Rm = (G << 1) | (gp >> 31);
        const uint32_t out = G & (Lm | Rm) & (yn[k] | yp[k]) & (col[k] | col[k + 2]);
        if (out != G) {
            bits[(size_t)z * planeW + col0] = out;
            col[k + 1] = out;  // later planes of this column see the update
            changed = true;
        }
    }
    // Flags, not counts: plain stores of 1 (atomics from every changing
    // wave to one address would serialise).
    const unsigned long long m = __ballot(changed);  // every lane takes part
    if (lane == 0 && m) {
        changes[pass * njobs + j] = 1u;
        totals[pass] = 1u;
        if (notconv) *notconv = 1u;
        stamp[L] = gpass + 1u;
    }
}

// Persistent pass: the grid is at most the resident workgroups; each XCD
// walks a contiguous run of logical blocks (its L2 then holds the neighbour
// planes).  After the first pass only the jobs the previous pass changed are
// visited (a batch may have one slow job among many finished ones).
__global__ __launch_bounds__(256) void k_erode(const Job *__restrict__ jobs,
                                               const uint32_t *__restrict__ erode_bases,
                                               uint32_t njobs, uint32_t y0, uint32_t y1,
                                               uint32_t z0, uint32_t z1,
                                               uint32_t *__restrict__ stamp,
                                               uint32_t *__restrict__ changes,
                                               uint32_t *__restrict__ totals,
                                               uint32_t *__restrict__ notconv, uint32_t pass,
                                               uint32_t gpass, int full) {
    const uint32_t total = erode_bases[njobs];  // bases[njobs] = all blocks
    const uint32_t nx = gridDim.x / kXcds;      // workgroups per XCD (grid: a multiple of 8)
    const uint32_t x = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const uint32_t per = (total + kXcds - 1u) / kXcds;
    const uint32_t lo = x * per, hi = min(lo + per, total);
    const uint32_t start = lo + slot;  // this workgroup takes start, start + nx, ...
    if (pass == 0) {
        for (uint32_t GL = start; GL < hi; GL += nx) {
            const uint32_t j = job_of(erode_bases, njobs, GL);
            if (erode_visit(jobs[j], GL, y0, y1, z0, z1, stamp, gpass, full))
                erode_block(jobs[j], j, GL, njobs, y0, y1, z0, z1, stamp, changes, totals,
                            notconv, pass, gpass);
        }
        return;
    }
    // Later passes: the workgroup's blocks are checked 256 at a time, one
    // per thread (its job changed in the previous pass, and the stamps near
    // it), and only the blocks that pass are visited, from a list in LDS.
    // (Checked one after another, each block would cost the workgroup a
    // round trip to memory, most of them to skip it.)
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_n;
    const uint32_t tid = threadIdx.x;
    const uint32_t *prev = changes + (pass - 1u) * njobs;
    for (uint32_t k0 = 0; start + k0 * nx < hi; k0 += blockDim.x) {  // uniform
        if (tid == 0) s_n = 0u;
        __syncthreads();
        const uint32_t GLt = start + (k0 + tid) * nx;
        if (GLt < hi) {
            const uint32_t jt = job_of(erode_bases, njobs, GLt);
            if (prev[jt] != 0u && erode_visit(jobs[jt], GLt, y0, y1, z0, z1, stamp, gpass, full))
                s_list[atomicAdd(&s_n, 1u)] = GLt;
        }
        __syncthreads();
        const uint32_t n = s_n;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t GL = s_list[i];
            const uint32_t j = job_of(erode_bases, njobs, GL);
            erode_block(jobs[j], j, GL, njobs, y0, y1, z0, z1, stamp, changes, totals, notconv,
                        pass, gpass);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Node writing
// ---------------------------------------------------------------------------
// Buffer offset past every range used below: loads there return 0 and
// stores there are dropped (raw buffer range check, offset >= size).
constexpr uint32_t kOob = 0x80000000u;

typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc rsrc(const VK_G void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t ld32(Rsrc r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ u32x4 ld128(Rsrc r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}
__device__ __forceinline__ void st128(Rsrc r, uint32_t off, u32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

// Nibble n (bit i = voxel i) -> 4 bytes holding 27 where the bit is set.
// The multiply moves bit i to bit 8i (its four partial products occupy
// disjoint bits, so nothing carries); 27 = 0b11011 is then two shifted ORs.
// The empty asm keeps them apart: folded together they become one 32-bit
// multiply by 27, which issues at a quarter of the rate of the two ORs.
__device__ __forceinline__ uint32_t nibble_27(uint32_t n) {
    const uint32_t s = __umul24(n, 0x204081u) & 0x01010101u;
    uint32_t t = s | (s << 1);
    __asm__("" : "+v"(t));
    return t | (t << 3);
}

// Cartesian node write: a block takes kChunk consecutive words of a plane
// (plane-linear, so no lane idles on narrow rows), each thread a column of
// kNodeZ planes.  All loads go through buffer resources, one per plane, so
// every address is a 32-bit lane offset and a neighbour outside the grid
// (z -1 / Z, rows -1 / Y, lanes past the plane) falls outside its range and
// reads 0 without a branch.  Every load of the column is issued before the
// first use, so a wave waits for memory once; the 32 material bytes of each
// surface word come in a second round trip.  x neighbours come from
// adjacent lanes.
//
// A word whose 32 voxels are all empty or all fully surrounded (the bulk of
// a grid) is stored as a constant.  Otherwise the fully surrounded voxels
// get id 27 by a nibble spread, boundary voxels (solid, some neighbour
// missing) get the id of their neighbour permutation from the bit-sliced
// table (orient_planes), patched byte-wise into the thread's LDS slot, and
// surface voxels OR in their material (bytes of the word's other voxels are
// 0 in the material grid).
//
// With several erosion subspaces (voxSplitRes y/z smaller than the grid) the
// reference computes a subspace's ids before later subspaces erode
// (src/host_side.cpp:1183-1206), so neighbours across a +y/+z seam are read
// from the pre-erosion snapshot `pre` (seam rows flagged in ystart/zstart);
// that path is a kernel of its own (k_nodes_cart_pre), so the common one
// holds no registers for it.
template <bool kPre>
__device__ __forceinline__ void nodes_cart(const Job *__restrict__ jobs,
                                           const uint32_t *__restrict__ node_bases,
                                           uint32_t njobs, const uint32_t *__restrict__ pre,
                                           const uint8_t *__restrict__ ystart,
                                           const uint8_t *__restrict__ zstart) {
    // Per wave 4 KB: the id bytes of each thread's word while boundary ids go
    // in (32 B per thread), then the wave's node run on its way out.
    __shared__ uint4 stage[256 * 4];
    const uint32_t GL = xcd_logical_block(node_bases[njobs]);  // bases[njobs] = all blocks
    if (GL == UINT32_MAX) return;  // uniform over the block
    const Job &J = jobs[job_of(node_bases, njobs, GL)];
    const GridDims g = J.gd;
    const int use_mats = J.use_mats, kind = J.kind;
    const bool zbrd = J.zero_border != 0, rot = J.rot_ids != 0;
    const BrickDims bd = J.bdn;
    const uint32_t P = g.W * g.Y;  // words per plane
    const uint32_t nch = (P + kChunk - 1u) / kChunk;
    const uint32_t lb = GL - J.node_base;  // block within the job
    const uint32_t tid = threadIdx.x, tx = tid & (kWave - 1u);
    const uint32_t zb = (lb / nch) * kNodeZ;
    const uint32_t i0 = (lb % nch) * kChunk, i = i0 + tid;
    const bool valid = i < P;
    const uint32_t y = valid ? i / g.W : 0u;
    const uint32_t w = valid ? i - y * g.W : 0u;
    const uint32_t P4 = P * 4u, W4 = g.W * 4u;
    // Byte offsets within a plane: the word, its y neighbours (row -1 wraps
    // past the range, row Y lands past it) and, for the wave's first and
    // last lane, the x neighbour held by the adjacent wave.
    const uint32_t o = valid ? i * 4u : kOob;
    const uint32_t oyn = valid ? o - W4 : kOob;
    const uint32_t oyp = o + W4;
    const uint32_t oex = valid && tx == 0u && w > 0u                 ? o - 4u
                         : valid && tx == kWave - 1u && w + 1u < g.W ? o + 4u
                                                                      : kOob;
    const VK_G uint32_t *bits = gptr(J.bits);
    auto plane = [&](const VK_G uint32_t *base, uint32_t z) {  // z -1 wraps to UINT32_MAX
        return z < g.Z ? rsrc(base + (size_t)z * P, P4) : rsrc(base, 0u);
    };
    // Round trip 1.
    uint32_t col[kNodeZ + 2], yn[kNodeZ], yp[kNodeZ], ex[kNodeZ];
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ + 2; k++) col[k] = ld32(plane(bits, zb + k - 1u), o);
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ; k++) {
        const Rsrc r = plane(bits, zb + k);
        yn[k] = ld32(r, oyn);
        yp[k] = ld32(r, oyp);
        ex[k] = ld32(r, oex);  // the x neighbour across the wave's edge (edge lanes)
    }
    uint32_t ypp[kNodeZ], zpp[kNodeZ];
    bool ys = false, zs[kNodeZ];
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ; k++) {
        ypp[k] = 0u;
        zpp[k] = 0u;
        zs[k] = false;
    }
    if (kPre) {
        const VK_G uint32_t *pg = gptr(pre);
        ys = valid && y + 1u < g.Y && ystart[y + 1u] != 0;
#pragma unroll
        for (uint32_t k = 0; k < kNodeZ; k++) {
            const uint32_t z = zb + k;
            ypp[k] = ld32(plane(pg, z), oyp);
            zpp[k] = ld32(plane(pg, z + 1u), o);
            zs[k] = z + 1u < g.Z && zstart[z + 1u] != 0;
        }
    }
    // Surface words (a word holding a surface voxel) from the brick's word
    // mask, and the offsets of their material bytes.
    bool S[kNodeZ];
    uint32_t moff[kNodeZ];
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ; k++) {
        S[k] = false;
        moff[k] = kOob;
    }
    Rsrc mr = rsrc(gptr(J.bits), 0u);
    if (use_mats) {  // uniform; the kNodeZ planes lie in one brick (kNodeZ divides its depth)
        // Material blocks relative to the first brick row the block touches.
        const uint32_t lbytes = bd.lwx + bd.lby + bd.lbz + 5u;  // log2 bytes per brick
        const uint32_t yb0 = (i0 / g.W) >> bd.lby;
        const uint32_t yb1 = ((min(i0 + kChunk, P) - 1u) / g.W) >> bd.lby;
        const uint32_t brow0 = ((zb >> bd.lbz) * bd.nby + yb0) * bd.nbx;
        const uint32_t nbrow = (yb1 - yb0 + 1u) * bd.nbx;  // bricks the block touches
        mr = rsrc(gptr(J.matgrid) + ((size_t)brow0 << lbytes), nbrow << lbytes);
        const uint32_t b = ((y >> bd.lby) - yb0) * bd.nbx + (w >> bd.lwx);
        const u32x4 hm = ld128(rsrc(gptr(J.wmask) + (size_t)brow0 * 4u, nbrow * 16u),
                               valid ? b * 16u : kOob);
        const uint32_t wib = word_in_brick(bd, w, y, zb);
        const uint32_t lane_off = (b << lbytes) + (wib << 5);
#pragma unroll
        for (uint32_t k = 0; k < kNodeZ; k++) {
            const uint32_t bit = wib + (k << (bd.lby + bd.lwx));
            const uint32_t q = bit >> 5;
            const uint32_t mw = q == 0u ? hm.x : q == 1u ? hm.y : q == 2u ? hm.z : hm.w;
            S[k] = ((mw >> (bit & 31u)) & 1u) != 0u;
            moff[k] = lane_off + ((k << (bd.lby + bd.lwx)) << 5);
        }
    }
    // Round trip 2: the 32 material bytes of each surface word.
    u32x4 ma[kNodeZ], mb[kNodeZ];
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ; k++) {
        ma[k] = u32x4{0u, 0u, 0u, 0u};
        mb[k] = ma[k];
        if (S[k]) {
            ma[k] = ld128(mr, moff[k]);
            mb[k] = ld128(mr, moff[k] + 16u);
        }
    }
    const uint32_t nbw = kind == 0 ? 32u : 64u;  // node bytes per word
    VK_G uint8_t *nodes = (VK_G uint8_t *)gptr(J.nodes);
    const uint32_t lane = tx;
    uint4 *wst = &stage[(tid - lane) * 4u];  // this wave's staging run
    uint8_t *sb = (uint8_t *)&wst[2u * lane];
    // The wave's 64 words are consecutive, so its nodes of a plane are one
    // run (2 KB ShortNode, 4 KB LongNode) starting at byte wrun: staged in
    // LDS in node order and stored lane by lane, each store writes 1 KB of
    // whole lines instead of every other 16 bytes.
    const uint32_t wrun = (i0 + (tid - lane)) * nbw;
#pragma unroll
    for (uint32_t k = 0; k < kNodeZ; k++) {
        const uint32_t z = zb + k;
        if (z >= g.Z) break;  // uniform
        const uint32_t G = col[k + 1];
        uint32_t gm = (uint32_t)__shfl_up((int)G, 1);
        uint32_t gp = (uint32_t)__shfl_down((int)G, 1);
        if (tx == 0) gm = ex[k];
        if (tx == kWave - 1) gp = ex[k];
        if (w == 0) gm = 0u;
        if (w + 1 >= g.W) gp = 0u;
        const uint32_t L = (G >> 1) | (gm << 31);
        const uint32_t R = (G << 1) | (gp >> 31);
        // Seam selects as masks: a select between two array elements
        // becomes a select between the arrays, which then live in scratch.
        const uint32_t ysm = ys ? ~0u : 0u, zsm = zs[k] ? ~0u : 0u;
        const uint32_t Yn = yn[k];
        const uint32_t Yp = (yp[k] & ~ysm) | (ypp[k] & ysm);
        const uint32_t Zn = col[k];
        const uint32_t Zp = (col[k + 2] & ~zsm) | (zpp[k] & zsm);
        const uint32_t interior = G & L & R & Yn & Yp & Zn & Zp;
        const uint32_t bnd = G & ~interior;  // solid, some neighbour missing
        uint32_t idw[8];
        if (bnd == 0u && !S[k]) {
            // All 32 voxels empty, or all fully surrounded (id 27).
            const uint32_t v = interior ? 0x1B1B1B1Bu : 0u;
#pragma unroll
            for (int j = 0; j < 8; j++) idw[j] = v;
        } else {
            const uint32_t ir = __brev(interior);  // bit i = voxel i
#pragma unroll
            for (int j = 0; j < 8; j++) idw[j] = nibble_27((ir >> (4 * j)) & 15u);
            if (bnd) {
                // A model rotated about z (x slices) takes its ids in the
                // un-rotated frame: its -x/+x are the rotated -y/+y, its -y/+y
                // the rotated +x/-x.
                const IdPlanes op = rot ? orient_planes(Yn, Yp, R, L, Zn, Zp)
                                        : orient_planes(L, R, Yn, Yp, Zn, Zp);
                wst[2u * lane] = make_uint4(idw[0], idw[1], idw[2], idw[3]);
                wst[2u * lane + 1u] = make_uint4(idw[4], idw[5], idw[6], idw[7]);
                uint32_t m = bnd;
                while (m) {
                    const uint32_t s = 31u - (uint32_t)__clz(m);  // voxel 31 - s
                    const uint32_t id = ((op.b[0] >> s) & 1u) | (((op.b[1] >> s) & 1u) << 1) |
                                        (((op.b[2] >> s) & 1u) << 2) |
                                        (((op.b[3] >> s) & 1u) << 3) |
                                        (((op.b[4] >> s) & 1u) << 4);
                    sb[31u - s] = (uint8_t)id;
                    m ^= 1u << s;
                }
                const uint4 q0 = wst[2u * lane], q1 = wst[2u * lane + 1u];
                idw[0] = q0.x; idw[1] = q0.y; idw[2] = q0.z; idw[3] = q0.w;
                idw[4] = q1.x; idw[5] = q1.y; idw[6] = q1.z; idw[7] = q1.w;
            }
        }
        const uint32_t md[8] = {ma[k].x, ma[k].y, ma[k].z, ma[k].w,
                                mb[k].x, mb[k].y, mb[k].z, mb[k].w};
        const Rsrc nr = rsrc(nodes + (size_t)z * P * nbw, P * nbw);
        // A partition with neighbours clears its border (zeroPadding,
        // src/voxelizer.cu:4811-4844): whole rows and planes, and the first
        // and last node of each row.
        const uint32_t all =
            zbrd && (y == 0u || y + 1u == g.Y || z == 0u || z + 1u == g.Z) ? 0u : ~0u;
        const bool x0 = zbrd && w == 0u, x1 = zbrd && w + 1u == g.W;
        if (kind == 0) {  // ShortNode byte: id | material << 5
            uint32_t out[8];
#pragma unroll
            for (int j = 0; j < 8; j++) out[j] = (idw[j] | ((md[j] << 5) & 0xE0E0E0E0u)) & all;
            if (x0) out[0] &= 0xFFFFFF00u;
            if (x1) out[7] &= 0x00FFFFFFu;
            wst[2u * lane] = make_uint4(out[0], out[1], out[2], out[3]);
            wst[2u * lane + 1u] = make_uint4(out[4], out[5], out[6], out[7]);
            wave_lds_sync();
#pragma unroll
            for (uint32_t q = 0; q < 2u; q++) {
                const uint4 v = wst[q * kWave + lane];
                st128(nr, wrun + (q * kWave + lane) * 16u, u32x4{v.x, v.y, v.z, v.w});
            }
        } else {  // LongNode halfword: id | material << 8
            uint32_t out[16];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                out[2 * j] = __builtin_amdgcn_perm(md[j], idw[j], 0x05010400u) & all;
                out[2 * j + 1] = __builtin_amdgcn_perm(md[j], idw[j], 0x07030602u) & all;
            }
            if (x0) out[0] &= 0xFFFF0000u;
            if (x1) out[15] &= 0x0000FFFFu;
#pragma unroll
            for (uint32_t q = 0; q < 4u; q++)
                wst[4u * lane + q] = make_uint4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
            wave_lds_sync();
#pragma unroll
            for (uint32_t q = 0; q < 4u; q++) {
                const uint4 v = wst[q * kWave + lane];
                st128(nr, wrun + (q * kWave + lane) * 16u, u32x4{v.x, v.y, v.z, v.w});
            }
        }
        wave_lds_sync();  // the run is read out before the next plane's writes
    }
}
__global__ __launch_bounds__(256) void k_nodes_cart(const Job *__restrict__ jobs,
                                                    const uint32_t *__restrict__ node_bases,
                                                    uint32_t njobs) {
    nodes_cart<false>(jobs, node_bases, njobs, nullptr, nullptr, nullptr);
}
// The seam path holds 8 more registers per thread; it runs at one wave per
// SIMD less rather than spill.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 5))) void k_nodes_cart_pre(
    const Job *__restrict__ jobs, const uint32_t *__restrict__ node_bases, uint32_t njobs,
    const uint32_t *__restrict__ pre, const uint8_t *__restrict__ ystart,
    const uint8_t *__restrict__ zstart) {
    nodes_cart<true>(jobs, node_bases, njobs, pre, ystart, zstart);
}

// PartialNode expansion (processVoxel<PartialNode>, src/voxelizer.cu:
// 2997-3033; r = float(solid) from constructNodeList2, 1675, which erosion
// leaves alone, so the bits are read from the pre-erosion copy): one thread
// per node.  A surface node's winner key comes from the keygrid at the node's
// position in its brick (as its material in matgrid); the ratio is computed for
// the winner at the node's alloc coordinates.  A cleared partition border
// keeps r = 0 (zeroPadding stores Node()).
__global__ __launch_bounds__(256) void k_nodes_partial(const Job *__restrict__ jobs, uint32_t j) {
    const Job &J = jobs[j];
    const GridDims g = J.gd;
    const BrickDims bd = J.bdn;
    const size_t n = (size_t)g.X * g.Y * g.Z;
    const VK_G uint16_t *ln = (const VK_G uint16_t *)gptr(J.nodes);
    VK_G u32x2 *out = (VK_G u32x2 *)J.pnodes;
    const VK_G uint32_t *bits = gptr(J.pbits);  // before erosion
    const bool zbrd = J.zero_border != 0, mats = J.use_mats != 0, embed = J.embed != 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(i % g.X), y = (uint32_t)((i / g.X) % g.Y);
        const uint32_t z = (uint32_t)(i / ((size_t)g.X * g.Y));
        const uint32_t w = x >> 5;
        const uint32_t bit =
            (bits[((size_t)z * g.Y + y) * g.W + w] >> (31u - (x & 31u))) & 1u;
        float r = bit ? 1.0f : 0.0f;
        if (mats && bit) {
            const uint32_t b = brick_of_word(bd, w, y, z);
            const size_t wi = ((size_t)z * g.Y + y) * g.W + w;
            if (gptr(J.bhist)[b] && ((gptr(J.surf)[wi] >> (31u - (x & 31u))) & 1u)) {
                const uint32_t lvox = bd.lwx + bd.lby + bd.lbz + 5u;  // log2 voxels per brick
                const uint32_t key = gptr(J.keygrid)[((size_t)b << lvox) +
                                                     (word_in_brick(bd, w, y, z) << 5) + (x & 31u)];
                const uint32_t t = embed ? (key >> 8) & 0x3FFFFFu : key & 0x3FFFFFFFu;
                const int32_t node[3] = {(int32_t)x, (int32_t)y, (int32_t)z};
                r = partial_ratio(J.V, J.I, t, key >> 30, node, J.m, J.fd);
            }
        }
        if (zbrd && (x == 0u || x + 1u == g.X || y == 0u || y + 1u == g.Y || z == 0u ||
                     z + 1u == g.Z))
            r = 0.0f;
        out[i] = u32x2{(uint32_t)ln[i], __float_as_uint(r)};
    }
}

// ---------------------------------------------------------------------------
// FCC nodes
// ---------------------------------------------------------------------------
// Spreads 16 bits to the even positions: bit j -> bit 2j.
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

struct FccGrids {
    const VK_G uint32_t *s[4];  // per-grid solid|surface bits (grid 1..4)
};

// 32 FCC nodes of row (y, fz), x in [32*fw, 32*fw+32): even x from the
// sub-grid with x shift 0, odd x from the one with x shift d/2
// (convertToFCCGrid, src/voxelizer.cu:1756-1772).
__device__ __forceinline__ uint32_t fcc_word(const FccGrids &G, const GridDims &g, int64_t fw,
                                             int64_t y, int64_t fz) {
    if (fw < 0 || fw >= (int64_t)(2 * g.W) || y < 0 || y >= (int64_t)g.Y || fz < 0 ||
        fz >= (int64_t)(2 * g.Z))
        return 0u;
    const int odd = (int)(fz & 1);
    // Picked by arithmetic: a select between elements of the pointer array
    // becomes a computed index, which puts the array in scratch.
    const uintptr_t e0 = (uintptr_t)G.s[0], o0 = (uintptr_t)G.s[1];
    const uintptr_t k = (uintptr_t)odd;
    const VK_G uint32_t *ge = (const VK_G uint32_t *)(e0 + k * ((uintptr_t)G.s[2] - e0));
    const VK_G uint32_t *go = (const VK_G uint32_t *)(o0 + k * ((uintptr_t)G.s[3] - o0));
    const size_t wi = ((size_t)(fz >> 1) * g.Y + (size_t)y) * g.W + (size_t)(fw >> 1);
    const int sh = (fw & 1) ? 0 : 16;
    const uint32_t e = (ge[wi] >> sh) & 0xFFFFu, o = (go[wi] >> sh) & 0xFFFFu;
    return (spread16(e) << 1) | spread16(o);
}

// Left/right shifted FCC row words: neighbours at x-1 / x+1 (the word of the
// adjacent lane supplies the carried bit).
__device__ __forceinline__ void lr_shift(const FccGrids &G, const GridDims &g, uint32_t word,
                                         int64_t fw, int64_t y, int64_t fz, uint32_t &Lw,
                                         uint32_t &Rw) {
    const uint32_t tx = threadIdx.x;
    uint32_t prev = (uint32_t)__shfl_up((int)word, 1);
    uint32_t next = (uint32_t)__shfl_down((int)word, 1);
    if (tx == 0) prev = fcc_word(G, g, fw - 1, y, fz);
    if (tx == kWave - 1) next = fcc_word(G, g, fw + 1, y, fz);
    Lw = (word >> 1) | (prev << 31);
    Rw = (word << 1) | (next >> 31);
}

// FCC node write, row-major like k_nodes_cart: the 12-bit neighbour mask of
// calculateFCCBoundaries (src/voxelizer.cu:1930-2048) as bitwise selects by
// the (x+z) parity mask, materials from the four sub-grids' surface lists.
__global__ __launch_bounds__(256) void k_nodes_fcc(const Job *__restrict__ jobs,
                                                   const uint32_t *__restrict__ fcc_first,
                                                   const uint32_t *__restrict__ fcc_bases,
                                                   uint32_t nfcc) {
    __shared__ uint4 fslot[256][4];  // 32 node masks (u16) per thread while boundary ids go in
    const uint32_t GL = xcd_logical_block(fcc_bases[nfcc]);  // bases[nfcc] = all blocks
    if (GL == UINT32_MAX) return;
    const uint32_t f = job_of(fcc_bases, nfcc, GL);
    const Job *__restrict__ Js = jobs + fcc_first[f];
    const GridDims g = Js[0].gd;
    const int use_mats = Js[0].use_mats, kind = Js[0].kind;
    VK_G void *__restrict__ nodes = gptr(Js[0].nodes);
    const BrickDims bd = Js[0].bdn;
    const VK_G uint32_t *s1 = gptr(Js[0].bits), *s2 = gptr(Js[1].bits), *s3 = gptr(Js[2].bits),
                        *s4 = gptr(Js[3].bits);
    const uint32_t FW = 2 * g.W, FY = g.Y;
    const uint32_t nxc = (FW + kWave - 1) / kWave, ny4 = (FY + 3) / 4;
    const uint32_t L = GL - fcc_bases[f];  // block within the lattice
    const RowBlock rb = row_block(L, nxc, ny4);
    const int64_t fw = (int64_t)rb.xc * kWave + threadIdx.x;
    const int64_t y = (int64_t)rb.y4 * 4 + threadIdx.y;
    const int64_t fz = rb.z;
    const bool valid = fw < (int64_t)FW && y < (int64_t)FY;
    FccGrids G{{s1, s2, s3, s4}};
    // All lanes take part in the shuffles; invalid lanes hold zero words.
    auto word = [&](int64_t yy, int64_t zz) { return valid ? fcc_word(G, g, fw, yy, zz) : 0u; };
    const uint32_t M0 = word(y, fz), Mm = word(y - 1, fz), Mp = word(y + 1, fz);
    const uint32_t D0 = word(y, fz - 1), Dm = word(y - 1, fz - 1), Dp = word(y + 1, fz - 1);
    const uint32_t U0 = word(y, fz + 1), Um = word(y - 1, fz + 1), Up = word(y + 1, fz + 1);
    uint32_t LD0, RD0, LMm, RMm, LM0, RM0, LMp, RMp, LU0, RU0;
    lr_shift(G, g, D0, fw, y, fz - 1, LD0, RD0);
    lr_shift(G, g, Mm, fw, y - 1, fz, LMm, RMm);
    lr_shift(G, g, M0, fw, y, fz, LM0, RM0);
    lr_shift(G, g, Mp, fw, y + 1, fz, LMp, RMp);
    lr_shift(G, g, U0, fw, y, fz + 1, LU0, RU0);
    if (!valid) return;
    const uint32_t E = (fz & 1) ? 0x55555555u : 0xAAAAAAAAu;  // (x+z) even
    const uint32_t NE = ~E;
    uint32_t n[12];
    n[0] = (E & Dm) | (NE & D0);
    n[1] = LD0;
    n[2] = RD0;
    n[3] = (E & D0) | (NE & Dp);
    n[4] = (E & LMm) | (NE & LM0);
    n[5] = (E & RMm) | (NE & RM0);
    n[6] = (E & LM0) | (NE & LMp);
    n[7] = (E & RM0) | (NE & RMp);
    n[8] = (E & Um) | (NE & U0);
    n[9] = LU0;
    n[10] = RU0;
    n[11] = (E & U0) | (NE & Up);
    uint32_t interior = M0;
#pragma unroll
    for (int j = 0; j < 12; j++) interior &= n[j];
    const uint32_t boundary = M0 & ~interior;
    // Neighbour masks, two nodes per dword (node 2d in the low half): 4095
    // for fully surrounded nodes, 0 for empty ones; boundary nodes get their
    // 12 neighbour bits, patched halfword-wise in the thread's LDS slot.
    const uint32_t ir = __brev(interior);  // bit k = node k
    uint32_t mk[16];
#pragma unroll
    for (int d = 0; d < 16; d++) {
        const uint32_t two = (ir >> (2 * d)) & 3u;
        mk[d] = ((two & 1u) ? 0x0FFFu : 0u) | ((two & 2u) ? 0x0FFF0000u : 0u);
    }
    if (boundary) {
        uint4 *my = fslot[threadIdx.y * kWave + threadIdx.x];
#pragma unroll
        for (int q = 0; q < 4; q++)
            my[q] = make_uint4(mk[4 * q], mk[4 * q + 1], mk[4 * q + 2], mk[4 * q + 3]);
        uint16_t *mh = (uint16_t *)my;
        uint32_t m = boundary;
        while (m) {
            const uint32_t s = 31u - (uint32_t)__clz(m);  // node 31 - s
            uint32_t v = 0u;
#pragma unroll
            for (int j = 0; j < 12; j++) v |= ((n[j] >> s) & 1u) << j;
            mh[31u - s] = (uint16_t)v;
            m ^= 1u << s;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 v = my[q];
            mk[4 * q] = v.x; mk[4 * q + 1] = v.y; mk[4 * q + 2] = v.z; mk[4 * q + 3] = v.w;
        }
    }
    // Materials: node k is voxel k / 2 of the sub-grid with x parity k & 1,
    // whose surface word's half for this FCC word holds 16 voxels.  A word
    // with a surface voxel has all 32 material bytes written in the call (0
    // for its other voxels), so each half is read as 16 bytes, or taken as 0
    // when it has no surface voxel.
    uint32_t ma[2][4] = {{0u, 0u, 0u, 0u}, {0u, 0u, 0u, 0u}};
    if (use_mats) {
        const int odd = (int)(fz & 1);
        const uint64_t gw = (uint64_t)(fw >> 1), gz = (uint64_t)(fz >> 1);
        const uint32_t b = (uint32_t)((((gz >> bd.lbz) * bd.nby + ((uint64_t)y >> bd.lby)) *
                                       bd.nbx) + (gw >> bd.lwx));
        const size_t bbase = ((size_t)b << (bd.lwx + bd.lby + bd.lbz + 5u)) +
                             (size_t)word_in_brick(bd, (uint32_t)gw, y, (uint32_t)gz) * 32u +
                             ((fw & 1) ? 16u : 0u);
        const size_t gwi = (gz * g.Y + (uint64_t)y) * g.W + gw;
        const int sh = (fw & 1) ? 0 : 16;
#pragma unroll
        for (int par = 0; par < 2; par++) {
            const MatSource src = mat_source(Js[odd ? 2 + par : par]);  // uniform
            if (!src.bhist[b]) continue;
            if (((src.surf[gwi] >> sh) & 0xFFFFu) == 0u) continue;
            const u32x4 v = *(const VK_G u32x4 *)(src.matgrid + bbase);
            ma[par][0] = v.x; ma[par][1] = v.y; ma[par][2] = v.z; ma[par][3] = v.w;
        }
    }
    const size_t node0 = (((size_t)fz * FY + (size_t)y) * FW + (size_t)fw) * 32u;
    if (kind == 3) {  // ShortFCCNode: mat<<12 | bid&4095
        uint32_t out[16];
#pragma unroll
        for (int d = 0; d < 16; d++) {
            const uint32_t a = (ma[0][d >> 2] >> (8 * (d & 3))) & 0xFu;
            const uint32_t c = (ma[1][d >> 2] >> (8 * (d & 3))) & 0xFu;
            out[d] = mk[d] | (a << 12) | (c << 28);
        }
        VK_G u32x4 *dst = (VK_G u32x4 *)((VK_G uint8_t *)gptr(nodes) + node0 * 2u);
#pragma unroll
        for (int q = 0; q < 4; q++)
            dst[q] = u32x4{out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]};
    } else {  // LongFCCNode: {u16 bid, u8 mat, u8 0}
        VK_G u32x4 *dst = (VK_G u32x4 *)((VK_G uint8_t *)gptr(nodes) + node0 * 4u);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            uint32_t o[4];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int d = 2 * q + e;
                const uint32_t a = (ma[0][d >> 2] >> (8 * (d & 3))) & 0xFFu;
                const uint32_t c = (ma[1][d >> 2] >> (8 * (d & 3))) & 0xFFu;
                o[2 * e] = (mk[d] & 0xFFFFu) | (a << 16);
                o[2 * e + 1] = (mk[d] >> 16) | (c << 16);
            }
            dst[q] = u32x4{o[0], o[1], o[2], o[3]};
        }
    }
}

}  // namespace vk
