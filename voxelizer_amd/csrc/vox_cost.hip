// vox_cost.hip -- learned cost model of the surface brick schedule (gfx950).
//
// The surface pass (k_surf_brick) takes its bricks from a list.  The plain
// schedule lists the non-empty bricks in index order; the learned schedule
// puts the bricks that a small MLP (16 -> 64 -> 64 -> 1, ReLU, FP32, on 16
// per-brick features) predicts to be expensive on a list taken first, so
// they start early and the persistent workgroups finish together.  Training targets are the
// per-brick durations the brick pass records (s_memtime) on seeded meshes.
// Both schedules produce identical output: the order only decides which
// workgroup processes a brick and when.
//
// Kernels: per-brick features, MLP forward, MLP gradient (per-block partial
// sums reduced in a fixed order, so training is deterministic), Adam, and the
// screening and heavy-list passes of the learned order.
#include "vox_kernels.h"

namespace vk {

// ---------------------------------------------------------------------------
// Features of one brick, from the (triangle, brick) pairs of its list and
// the per-triangle class, brick box and voxel extents of k_surf_setup:
//  0 log1p(triangles)            1-3 log1p(1D / 2D / 3D triangles)
//  4-6 log1p(estimated items of 1D / 2D / 3D triangles in the brick)
//  7 log1p(estimated 3D voxel tests)   8 log1p(largest estimated items)
//  9 log1p(64-triangle chunks)   10-11 share of 3D triangles with axis x / y
// 12 log1p(bricks spanned, summed)  13 estimated items per brick voxel
// 14 log1p(triangles with <= 4 items)  15 constant 1
// A triangle's items in the brick are estimated as its items spread evenly
// over the bricks its box spans across its dominant plane.
// ---------------------------------------------------------------------------
// Wave reductions in registers (DPP; every lane active): within each 16-lane
// row by the quad swaps and the half-row and row mirrors, then the rows into
// lane 63 by the row broadcasts (rows 1 and 3 take rows 0 and 2, rows 2 and 3
// take lane 31), read back as a uniform value.  Masked-off rows keep `old`.
template <int kCtrl, int kRows>
__device__ __forceinline__ float dpp_f32(float old, float x) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), kCtrl, kRows, 0xF, false));
}
template <class Op>
__device__ __forceinline__ float wave_reduce(float v, float identity, Op op) {
    v = op(v, dpp_f32<0xB1, 0xF>(identity, v));   // quad_perm [1,0,3,2]
    v = op(v, dpp_f32<0x4E, 0xF>(identity, v));   // quad_perm [2,3,0,1]
    v = op(v, dpp_f32<0x141, 0xF>(identity, v));  // row_half_mirror
    v = op(v, dpp_f32<0x140, 0xF>(identity, v));  // row_mirror
    v = op(v, dpp_f32<0x142, 0xA>(identity, v));  // row_bcast:15
    v = op(v, dpp_f32<0x143, 0xC>(identity, v));  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), kWave - 1));
}
__device__ __forceinline__ float wave_sum(float v) {
    return wave_reduce(v, 0.0f, [](float a, float b) { return a + b; });
}
// Maxima of non-negative values: the identity 0 never decides.
__device__ __forceinline__ float wave_max(float v) {
    return wave_reduce(v, 0.0f, [](float a, float b) { return fmaxf(a, b); });
}

// Per-lane sums of the features (feature order: see above).
struct FeatAcc {
    float c1, c2, c3, i1, i2, i3, v3, mx, span, small, axx, axy;
};

// Accumulates triangles first, first + stride, ... of brick b's list (each
// thread in index order).  Returns the brick's triangle count and voxels.
__device__ __forceinline__ void feat_accumulate(uint32_t b, const uint32_t *__restrict__ brick_hist,
                                                const uint32_t *__restrict__ brick_off,
                                                const uint32_t *__restrict__ blist,
                                                const uint4 *__restrict__ tri_info,
                                                const uint4 *__restrict__ tri_ext,
                                                const Job *__restrict__ jobs,
                                                const uint32_t *__restrict__ brick_bases,
                                                uint32_t njobs, uint32_t first, uint32_t stride,
                                                FeatAcc &A, uint32_t &cnt, float &bvox) {
    const Job &J = jobs[job_of(brick_bases, njobs, b)];
    const int32_t bd0 = J.bdim[0], bd1 = J.bdim[1], bd2 = J.bdim[2];
    bvox = (float)(bd0 * bd1 * bd2);
    cnt = brick_hist[b];
    const uint32_t off = brick_off[b];
    // Eight triangles per thread per step, their loads issued together (a
    // heavy brick's hundreds of triangles would otherwise cost one round
    // trip per step).
    constexpr uint32_t kU = 8;
    for (uint32_t i0 = first; i0 < cnt; i0 += kU * stride) {
        uint32_t tt[kU];
        uint4 tis[kU], tes[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const uint32_t i = i0 + u * stride;
            tt[u] = i < cnt ? blist[off + i] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            const bool ok = i0 + u * stride < cnt;
            tis[u] = ok ? tri_info[tt[u]] : make_uint4(0u, 0u, 0u, 0u);
            tes[u] = ok ? tri_ext[tt[u]] : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; u++) {
            if (i0 + u * stride >= cnt) break;
            const uint4 ti = tis[u], te = tes[u];
            const int bb = (int)(ti.x >> 30), a = (int)((ti.x >> 28) & 3u);
            const uint32_t wbx = ti.z >> 16, wby = ti.w & 0xFFFFu, wbz = ti.w >> 16;
            const uint32_t wa = a == AX_X ? wbx : a == AX_Y ? wby : wbz;
            const uint32_t wuv = (wbx * wby * wbz) / max(wa, 1u);
            const int32_t bda = a == AX_X ? bd0 : a == AX_Y ? bd1 : bd2;
            float est;
            if (bb == BB_1D) {
                est = (float)te.w / (float)max(wa, 1u);
                A.c1 += 1.0f;
                A.i1 += est;
            } else {
                est = (float)te.w / (float)max(wuv, 1u);
                if (bb == BB_2D) {
                    A.c2 += 1.0f;
                    A.i2 += est;
                } else {
                    A.c3 += 1.0f;
                    A.i3 += est;
                    A.v3 += est * (float)min((int32_t)te.z, bda);
                    A.axx += a == AX_X ? 1.0f : 0.0f;
                    A.axy += a == AX_Y ? 1.0f : 0.0f;
                }
            }
            A.mx = fmaxf(A.mx, est);
            A.span += (float)(wbx * wby * wbz);
            A.small += est <= 4.0f ? 1.0f : 0.0f;
        }
    }
}

// Sums over the wave (butterflies: every lane ends with the totals).
__device__ __forceinline__ void feat_wave_reduce(FeatAcc &A) {
    A.c1 = wave_sum(A.c1); A.c2 = wave_sum(A.c2); A.c3 = wave_sum(A.c3);
    A.i1 = wave_sum(A.i1); A.i2 = wave_sum(A.i2); A.i3 = wave_sum(A.i3);
    A.v3 = wave_sum(A.v3); A.mx = wave_max(A.mx); A.span = wave_sum(A.span);
    A.small = wave_sum(A.small); A.axx = wave_sum(A.axx); A.axy = wave_sum(A.axy);
}

__device__ __forceinline__ void feat_finish(uint32_t cnt, float bvox, const FeatAcc &A,
                                            float f[kCostFeatures]) {
    f[0] = log1pf((float)cnt);
    f[1] = log1pf(A.c1);
    f[2] = log1pf(A.c2);
    f[3] = log1pf(A.c3);
    f[4] = log1pf(A.i1);
    f[5] = log1pf(A.i2);
    f[6] = log1pf(A.i3);
    f[7] = log1pf(A.v3);
    f[8] = log1pf(A.mx);
    f[9] = log1pf((float)((cnt + 63u) / 64u));
    f[10] = A.c3 > 0.0f ? A.axx / A.c3 : 0.0f;
    f[11] = A.c3 > 0.0f ? A.axy / A.c3 : 0.0f;
    f[12] = log1pf(A.span);
    f[13] = (A.i1 + A.i2 + A.i3) / bvox;
    f[14] = log1pf(A.small);
    f[15] = 1.0f;
}

// The 16 features of brick b, computed by one wave (every lane ends with
// all of them).
__device__ __forceinline__ void brick_feats(uint32_t b, const uint32_t *__restrict__ brick_hist,
                                            const uint32_t *__restrict__ brick_off,
                                            const uint32_t *__restrict__ blist,
                                            const uint4 *__restrict__ tri_info,
                                            const uint4 *__restrict__ tri_ext,
                                            const Job *__restrict__ jobs,
                                            const uint32_t *__restrict__ brick_bases,
                                            uint32_t njobs, float f[kCostFeatures]) {
    FeatAcc A = {};
    uint32_t cnt;
    float bvox;
    feat_accumulate(b, brick_hist, brick_off, blist, tri_info, tri_ext, jobs, brick_bases, njobs,
                    lane_id(), kWave, A, cnt, bvox);
    feat_wave_reduce(A);
    feat_finish(cnt, bvox, A, f);
}

// Row r describes brick order[sel ? sel[r] : r], for r < *n_rows.
__global__ __launch_bounds__(256) void k_brick_features(
    const uint32_t *__restrict__ order, const uint32_t *__restrict__ sel,
    const uint32_t *__restrict__ n_rows, const uint32_t *__restrict__ brick_hist,
    const uint32_t *__restrict__ brick_off, const uint32_t *__restrict__ blist,
    const uint4 *__restrict__ tri_info, const uint4 *__restrict__ tri_ext,
    const Job *__restrict__ jobs, const uint32_t *__restrict__ brick_bases, uint32_t njobs,
    float *__restrict__ feats) {
    const uint32_t n = *n_rows;
    const uint32_t nwaves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; q < n; q += nwaves) {
        float f[kCostFeatures];
        brick_feats(order[sel ? sel[q] : q], brick_hist, brick_off, blist, tri_info, tri_ext,
                    jobs, brick_bases, njobs, f);
        if (lane_id() == 0) {
            float *o = feats + (size_t)q * kCostFeatures;
#pragma unroll
            for (int i = 0; i < kCostFeatures; i++) o[i] = f[i];
        }
    }
}

// ---------------------------------------------------------------------------
// MLP.  Parameter layout (kCostParams = 5313 floats): W1[64][16] at 0,
// b1[64] at 1024, W2[64][64] at 1088, b2[64] at 5184, W3[64] at 5248, b3 at
// 5312.  A workgroup of 256 threads takes kMlpRows samples: lane j is hidden
// neuron j, the four waves split the samples.  Weights sit in LDS padded to
// odd row strides (17, 65 floats) so both row and column walks are free of
// bank conflicts.  Sums run in a fixed order (inputs in index order, the
// output layer as a wave butterfly).
// ---------------------------------------------------------------------------
constexpr int kIn = kCostFeatures, kHid = 64, kMlpRows = (int)kMlpBlockRows;
constexpr int kOffB1 = 1024, kOffW2 = 1088, kOffB2 = 5184, kOffW3 = 5248, kOffB3 = 5312;

struct MlpShared {
    float w1[kHid][kIn + 1];
    float w2[kHid][kHid + 1];
    float b1[kHid], b2[kHid], w3[kHid];
    float b3;
    alignas(16) float x[kMlpRows][kIn];
    alignas(16) float h1[kMlpRows][kHid];
    alignas(16) float h2[kMlpRows][kHid];
};

__device__ void mlp_load(MlpShared &S, const float *__restrict__ P, const float *__restrict__ X,
                         uint32_t row0, uint32_t n) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < kHid * kIn; k += blockDim.x) S.w1[k / kIn][k % kIn] = P[k];
    for (uint32_t k = tid; k < kHid * kHid; k += blockDim.x)
        S.w2[k / kHid][k % kHid] = P[kOffW2 + k];
    if (tid < (uint32_t)kHid) {
        S.b1[tid] = P[kOffB1 + tid];
        S.b2[tid] = P[kOffB2 + tid];
        S.w3[tid] = P[kOffW3 + tid];
    }
    if (tid == 0) S.b3 = P[kOffB3];
    for (uint32_t k = tid; k < kMlpRows * kIn; k += blockDim.x) {
        const uint32_t r = row0 + k / kIn;
        S.x[k / kIn][k % kIn] = r < n ? X[(size_t)r * kIn + k % kIn] : 0.0f;
    }
    __syncthreads();
}

// Hidden layers for the block's rows; returns the outputs of rows
// wave, wave+4, ... in y[] (valid in every lane).
// Neuron j's weight rows live in registers; the row inputs are read as
// 16-byte LDS broadcasts.
__device__ __forceinline__ void mlp_hidden(MlpShared &S, float y[kMlpRows / 4]) {
    const uint32_t j = lane_id(), w = threadIdx.x / kWave;
    float r1[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) r1[i] = S.w1[j][i];
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int s = (int)w + 4 * k;
        float z = S.b1[j];
        const float4 *xv = (const float4 *)S.x[s];
#pragma unroll
        for (int i = 0; i < kIn / 4; i++) {
            const float4 v = xv[i];
            z += r1[4 * i] * v.x;
            z += r1[4 * i + 1] * v.y;
            z += r1[4 * i + 2] * v.z;
            z += r1[4 * i + 3] * v.w;
        }
        S.h1[s][j] = fmaxf(z, 0.0f);
    }
    __syncthreads();
    float r2[kHid];
#pragma unroll
    for (int i = 0; i < kHid; i++) r2[i] = S.w2[j][i];
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int s = (int)w + 4 * k;
        float z = S.b2[j];
        const float4 *hv = (const float4 *)S.h1[s];
#pragma unroll
        for (int i = 0; i < kHid / 4; i++) {
            const float4 v = hv[i];
            z += r2[4 * i] * v.x;
            z += r2[4 * i + 1] * v.y;
            z += r2[4 * i + 2] * v.z;
            z += r2[4 * i + 3] * v.w;
        }
        const float h = fmaxf(z, 0.0f);
        S.h2[s][j] = h;
        y[k] = wave_sum(S.w3[j] * h) + S.b3;
    }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ P,
                                                     const float *__restrict__ X, uint32_t n,
                                                     float *__restrict__ Y) {
    __shared__ MlpShared S;
    const uint32_t row0 = blockIdx.x * kMlpRows;
    mlp_load(S, P, X, row0, n);
    float y[kMlpRows / 4];
    mlp_hidden(S, y);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int k = 0; k < kMlpRows / 4; k++) {
            const uint32_t r = row0 + w + 4 * k;
            if (r < n) Y[r] = y[k];
        }
}

// Gradient of the mean squared error over rows [0, n) (scaled by 1/n_scale)
// for the block's rows: the block writes its partial sums to
// partial[block][kCostParams] and its squared-error sum to ploss[block].
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ P,
                                                  const float *__restrict__ X,
                                                  const float *__restrict__ T, uint32_t n,
                                                  float inv_scale, float *__restrict__ partial,
                                                  float *__restrict__ ploss) {
    __shared__ MlpShared S;
    __shared__ float dz2[kMlpRows][kHid];
    __shared__ float dz1[kMlpRows][kHid];
    __shared__ float dy[kMlpRows];
    __shared__ float lsum[4];
    const uint32_t row0 = blockIdx.x * kMlpRows;
    mlp_load(S, P, X, row0, n);
    float y[kMlpRows / 4];
    mlp_hidden(S, y);
    const uint32_t j = lane_id(), w = threadIdx.x / kWave;
    float l = 0.0f;
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int s = (int)w + 4 * k;
        const uint32_t r = row0 + (uint32_t)s;
        const float e = r < n ? y[k] - T[r] : 0.0f;
        l += e * e;
        const float g = 2.0f * e * inv_scale;
        if (j == 0) dy[s] = g;
        dz2[s][j] = S.h2[s][j] > 0.0f ? g * S.w3[j] : 0.0f;
    }
    if (j == 0) lsum[w] = l;
    __syncthreads();
    for (int k = 0; k < kMlpRows / 4; k++) {
        const int s = (int)w + 4 * k;
        float d = 0.0f;  // dL/dh1[s][j] = sum_k W2[k][j] dz2[s][k]
        for (int q = 0; q < kHid; q++) d += S.w2[q][j] * dz2[s][q];
        dz1[s][j] = S.h1[s][j] > 0.0f ? d : 0.0f;
    }
    __syncthreads();
    float *out = partial + (size_t)blockIdx.x * kCostParams;
    // dW1[j][i], i = 4w .. 4w+3
    for (int ii = 0; ii < kIn / 4; ii++) {
        const int i = (int)w * (kIn / 4) + ii;
        float g = 0.0f;
        for (int s = 0; s < kMlpRows; s++) g += dz1[s][j] * S.x[s][i];
        out[j * kIn + i] = g;
    }
    // dW2[j][i], i = 16w .. 16w+15
    for (int ii = 0; ii < kHid / 4; ii++) {
        const int i = (int)w * (kHid / 4) + ii;
        float g = 0.0f;
        for (int s = 0; s < kMlpRows; s++) g += dz2[s][j] * S.h1[s][i];
        out[kOffW2 + j * kHid + i] = g;
    }
    float g = 0.0f;
    if (w == 0) {
        for (int s = 0; s < kMlpRows; s++) g += dz1[s][j];
        out[kOffB1 + j] = g;
    } else if (w == 1) {
        for (int s = 0; s < kMlpRows; s++) g += dz2[s][j];
        out[kOffB2 + j] = g;
    } else if (w == 2) {
        for (int s = 0; s < kMlpRows; s++) g += dy[s] * S.h2[s][j];
        out[kOffW3 + j] = g;
    } else if (j == 0) {
        for (int s = 0; s < kMlpRows; s++) g += dy[s];
        out[kOffB3] = g;
        ploss[blockIdx.x] = (lsum[0] + lsum[1]) + (lsum[2] + lsum[3]);
    }
}

// grad[k] = sum over blocks b (in order) of partial[b][k]; loss likewise.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float *__restrict__ partial,
                                                    const float *__restrict__ ploss,
                                                    uint32_t nblocks, float *__restrict__ grad,
                                                    float *__restrict__ loss) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < (uint32_t)kCostParams) {
        float g = 0.0f;
        for (uint32_t b = 0; b < nblocks; b++) g += partial[(size_t)b * kCostParams + k];
        grad[k] = g;
    } else if (k == (uint32_t)kCostParams && loss) {
        float l = 0.0f;
        for (uint32_t b = 0; b < nblocks; b++) l += ploss[b];
        *loss = l;
    }
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, const float *__restrict__ G,
                                              float *__restrict__ M, float *__restrict__ V,
                                              uint32_t n, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float g = G[k];
    const float m = b1 * M[k] + (1.0f - b1) * g;
    const float v = b2 * V[k] + (1.0f - b2) * g * g;
    M[k] = m;
    V[k] = v;
    P[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// ---------------------------------------------------------------------------
// Learned brick order.  Candidates are the listed bricks whose pair count is
// at least `ratio` times the mean (the plain counts screen the list); the MLP
// predicts their cost, and those predicted heavy (standardized log cost above
// `heavy`) go on a short list that the brick pass works through first, one
// entry per grab.  A brick predicted more than kSplitStep above `heavy` is cut
// into 2 sub-bricks along z, more than 2 kSplitStep into 4 (at most its
// depth and max_parts): each entry carries its layer range (hsub: first | end << 16), so
// the workgroups that take them share the brick's work (voxels of different
// layers are disjoint, so the split does not change the winners).  The
// plain list follows in index order (neighbouring bricks share triangles and
// solid rows in L2), skipping the bricks already taken.  cnt[0] counts
// candidates, cnt[1] heavy entries; both start at zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cost_screen(const uint32_t *__restrict__ order,
                                                     const uint32_t *__restrict__ n_order,
                                                     const uint32_t *__restrict__ brick_hist,
                                                     const uint32_t *__restrict__ total_pairs,
                                                     float ratio, uint32_t *__restrict__ cand,
                                                     uint32_t *__restrict__ cnt,
                                                     uint8_t *__restrict__ skip) {
    const uint32_t n = *n_order;
    const float thr = n ? ratio * (float)(*total_pairs) / (float)n : 0.0f;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const bool c = (float)brick_hist[order[q]] >= thr;
        skip[q] = 0u;
        const uint32_t slot = wave_bump(&cnt[0], c);  // one atomic per wave
        if (c) cand[slot] = q;
    }
}

// The learned path in one launch: for each candidate (one workgroup per
// brick, its four waves taking every fourth run of 64 triangles, their sums
// combined in a fixed order) its features, the MLP forward pass on wave 0
// with lane j as hidden neuron j (weights read from global memory, where they
// stay cached; the sums run in the order of k_mlp_forward), and the heavy
// test that puts the brick on the heavy list and flags its plain entry.
__global__ __launch_bounds__(256) void k_cost_predict(
    const float *__restrict__ P, const uint32_t *__restrict__ order,
    const uint32_t *__restrict__ cand, uint32_t *__restrict__ cnt, float heavy,
    uint32_t max_parts, const uint32_t *__restrict__ brick_hist,
    const uint32_t *__restrict__ brick_off,
    const uint32_t *__restrict__ blist, const uint4 *__restrict__ tri_info,
    const uint4 *__restrict__ tri_ext, const Job *__restrict__ jobs,
    const uint32_t *__restrict__ brick_bases, uint32_t njobs, uint32_t *__restrict__ hlist,
    uint32_t *__restrict__ hsub, uint8_t *__restrict__ skip) {
    constexpr int kWaves = 256 / kWave;
    __shared__ FeatAcc part[kWaves];
    const uint32_t n = cnt[0];
    const uint32_t j = lane_id(), w = threadIdx.x / kWave;
    for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {  // uniform over the block
        const uint32_t q = cand[r], b = order[q];
        FeatAcc A = {};
        uint32_t tris;
        float bvox;
        feat_accumulate(b, brick_hist, brick_off, blist, tri_info, tri_ext, jobs, brick_bases,
                        njobs, threadIdx.x, blockDim.x, A, tris, bvox);
        feat_wave_reduce(A);
        if (j == 0) part[w] = A;
        __syncthreads();
        if (w == 0) {
            const FeatAcc p0 = part[0], p1 = part[1], p2 = part[2], p3 = part[3];
            auto add = [](float a, float b, float c, float d) { return (a + b) + (c + d); };
            A.c1 = add(p0.c1, p1.c1, p2.c1, p3.c1);
            A.c2 = add(p0.c2, p1.c2, p2.c2, p3.c2);
            A.c3 = add(p0.c3, p1.c3, p2.c3, p3.c3);
            A.i1 = add(p0.i1, p1.i1, p2.i1, p3.i1);
            A.i2 = add(p0.i2, p1.i2, p2.i2, p3.i2);
            A.i3 = add(p0.i3, p1.i3, p2.i3, p3.i3);
            A.v3 = add(p0.v3, p1.v3, p2.v3, p3.v3);
            A.mx = fmaxf(fmaxf(p0.mx, p1.mx), fmaxf(p2.mx, p3.mx));
            A.span = add(p0.span, p1.span, p2.span, p3.span);
            A.small = add(p0.small, p1.small, p2.small, p3.small);
            A.axx = add(p0.axx, p1.axx, p2.axx, p3.axx);
            A.axy = add(p0.axy, p1.axy, p2.axy, p3.axy);
            float f[kCostFeatures];
            feat_finish(tris, bvox, A, f);
            float z = P[kOffB1 + j];
#pragma unroll
            for (int i = 0; i < kIn; i++) z += P[j * kIn + i] * f[i];
            const float h1 = fmaxf(z, 0.0f);
            z = P[kOffB2 + j];
#pragma unroll
            for (int k = 0; k < kHid; k++)
                z += P[kOffW2 + j * kHid + k] *
                     __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), k));
            const float h2 = fmaxf(z, 0.0f);
            const float y = wave_sum(P[kOffW3 + j] * h2) + P[kOffB3];
            if (j == 0 && y > heavy) {  // false for NaN
                const uint32_t depth = (uint32_t)jobs[job_of(brick_bases, njobs, b)].bdim[2];
                uint32_t parts =
                    y > heavy + 2.0f * kSplitStep ? 4u : y > heavy + kSplitStep ? 2u : 1u;
                parts = min(min(parts, max_parts), depth);  // all powers of two
                const uint32_t base = atomicAdd(&cnt[1], parts);
                for (uint32_t k = 0; k < parts; k++) {
                    hlist[base + k] = b;
                    hsub[base + k] = (k * depth / parts) | (((k + 1u) * depth / parts) << 16);
                }
                skip[q] = 1u;
            }
        }
        __syncthreads();  // part[] is read before the next candidate writes it
    }
}

}  // namespace vk
