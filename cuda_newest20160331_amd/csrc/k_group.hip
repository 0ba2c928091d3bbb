// k_group.hip -- group averaging of correlation matrices and the dense
// threshold that turns the average into a CSR graph.
//
// The reference's group mode (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:418-670)
// adds every subject's r (or its Fisher z, :36-42) into a packed upper
// triangle of M1 floats (post_block, :824-887), divides by the subject count and
// maps Fisher sums back (:548-558), then thresholds the averages (FormFullGraph
// :683-703 or the exact-k FormFullGraph_s :743-822) and lists each row's
// neighbours in column order (FormCSRGraph :889-918).  Here the sums stay in
// HBM (the tile kernel's accumulation pass adds each subject), and the
// threshold writes a full N x N keep matrix of bytes (2 GB at N = 45000; rows
// padded to a multiple of 64 bytes), so
// that the symmetric CSR is read row by row with coalesced loads instead of
// gathering column entries from the packed triangle.
#include "common.hpp"

namespace bna {

typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int64_t packed_at(int64_t N, int64_t i, int64_t j)
{
    return i * N - (i + 1) * i / 2 + j - i - 1;
}

// Order-preserving 32-bit key of a float (larger value, larger key); -0 is
// taken as +0 so that the two compare equal, as floats do.
__device__ __forceinline__ uint32_t fkey(float v)
{
    uint32_t u = __builtin_bit_cast(uint32_t, v == 0.0f ? 0.0f : v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// CorMat.cpp:548-558: Cormat[i] /= FileNumber in float, then for Fisher sums
// inv_fishertrans (:44-50): r = exp(2 z), (r - 1) / (r + 1), all in float.
__global__ __launch_bounds__(256) void group_finish_kernel(float *__restrict__ acc, int64_t M1, float S, int fisher)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < M1; t += stride) {
        float v = acc[t] / S;
        if (fisher) {
            const float e = expf(2.0f * v);
            v = (e - 1.0f) / (e + 1.0f);
        }
        acc[t] = v;
    }
}

__global__ __launch_bounds__(256) void group_digits_kernel(const float *__restrict__ acc, int64_t M1, uint32_t prefix,
                                                           uint32_t mask, int shift,
                                                           unsigned long long *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // a run of equal digits (the common case for the top digit, where most
    // averages share an exponent) costs one LDS atomic
    uint32_t cur = 0, run = 0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < M1; t += stride) {
        const uint32_t k = fkey(acc[t]);
        if ((k & mask) != prefix) continue;
        const uint32_t d = (k >> shift) & 255u;
        if (d != cur && run) {
            atomicAdd(&h[cur], run);
            run = 0;
        }
        cur = d;
        run++;
    }
    if (run) atomicAdd(&h[cur], run);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// One wave per packed row.
__global__ __launch_bounds__(256) void group_ties_kernel(const float *__restrict__ acc, int N, float val,
                                                         int64_t *__restrict__ tie_rows)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N) return;
    int64_t c = 0;
    if (i + 1 < N) {
        const float *row = acc + packed_at(N, i, i + 1);  // columns i + 1 .. N - 1
        for (int64_t j = lane; j < N - 1 - i; j += 64) c += row[j] == val;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) tie_rows[i] = c;
}

// One 64 x 64 block (bi <= bj) of the upper triangle per workgroup: decide
// each pair, write the block to its place in the keep matrix and its
// transpose through LDS.  Rows of the keep matrix are ld bytes apart, ld a
// multiple of 64, so every thread's 16 bytes of a row go out as one store.
__global__ __launch_bounds__(256) void group_keep_kernel(const float *__restrict__ acc, int N, int64_t ld, float thr,
                                                         int dens, int64_t p_cut, uint8_t *__restrict__ keep,
                                                         unsigned int *__restrict__ flag)
{
    const int bj = blockIdx.x, bi = blockIdx.y;
    if (bi > bj) return;
    __shared__ __align__(16) uint8_t tile[64][80];  // row pitch: a multiple of 16 bytes
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)bi * 64, j0 = (int64_t)bj * 64;
    {
        const int r = t >> 2, cs = (t & 3) * 16;
        const int64_t i = i0 + r;
        bool bad = false;
        uint8_t k[16];
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const int64_t j = j0 + cs + c;
            k[c] = 0;
            if (i < N && j < N && i < j) {
                const int64_t p = packed_at(N, i, j);
                const float v = acc[p];
                k[c] = dens ? (uint8_t)(v > thr || (v == thr && p <= p_cut)) : (uint8_t)(v > thr);
                bad |= v > 1.0f;
            }
        }
        u32x4_t w;
#pragma unroll
        for (int q = 0; q < 4; q++)
            w[q] = (uint32_t)k[4 * q] | ((uint32_t)k[4 * q + 1] << 8) | ((uint32_t)k[4 * q + 2] << 16) |
                   ((uint32_t)k[4 * q + 3] << 24);
        *(u32x4_t *)&tile[r][cs] = w;
        // the block's own rows: pairs (i, j) with i <= j; below the diagonal of
        // a diagonal block the transpose below writes the same bytes' partners,
        // and here those positions hold 0 until it does
        if (i < N && bi < bj) *(u32x4_t *)(keep + i * ld + j0 + cs) = w;
        if (bad && !dens) atomicOr(flag, 1u);
    }
    __syncthreads();
    {
        const int jj = t >> 2, is = (t & 3) * 16;
        const int64_t j = j0 + jj;
        if (j < N) {
            uint8_t k[16];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                const int ii = is + c;
                // diagonal block: (j, i) with i < j comes from tile[ii][jj]; i >= j from tile[jj][ii]
                k[c] = (bi < bj || ii < jj) ? tile[ii][jj] : tile[jj][ii];
            }
            u32x4_t w;
#pragma unroll
            for (int q = 0; q < 4; q++)
                w[q] = (uint32_t)k[4 * q] | ((uint32_t)k[4 * q + 1] << 8) | ((uint32_t)k[4 * q + 2] << 16) |
                       ((uint32_t)k[4 * q + 3] << 24);
            *(u32x4_t *)(keep + j * ld + i0 + is) = w;
        }
    }
}

// One wave per row of the keep matrix: 16 bytes per lane per step.
__global__ __launch_bounds__(256) void group_rowcount_kernel(const uint8_t *__restrict__ keep, int N, int64_t ld,
                                                             int64_t *__restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N) return;
    const uint8_t *row = keep + i * ld;
    int64_t c = 0;
    for (int64_t j = (int64_t)lane * 16; j < N; j += 1024) {
        const u32x4_t w = *(const u32x4_t *)(row + j);
        // bytes are 0 or 1; columns at or past N are 0
#pragma unroll
        for (int q = 0; q < 4; q++) c += (int64_t)((w[q] * 0x01010101u) >> 24);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) cnt[i] = c;
}

__global__ __launch_bounds__(256) void group_fill_kernel(const uint8_t *__restrict__ keep, int64_t ld,
                                                         const float *__restrict__ acc, int N,
                                                         const int32_t *__restrict__ R, int32_t *__restrict__ C,
                                                         float *__restrict__ V)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N) return;
    const uint8_t *row = keep + i * ld;
    int64_t base = R[i];
    for (int64_t j0 = 0; j0 < N; j0 += 1024) {
        const int64_t jl = j0 + (int64_t)lane * 16;
        u32x4_t w = {0u, 0u, 0u, 0u};
        if (jl < N) w = *(const u32x4_t *)(row + jl);
        int n = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) n += (int)((w[q] * 0x01010101u) >> 24);
        // this lane's first slot: the kept columns of the lanes before it
        int pre = n;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(pre, d, 64);
            if (lane >= d) pre += o;
        }
        const int total = __shfl(pre, 63, 64);
        int64_t pos = base + pre - n;
        if (n) {
#pragma unroll
            for (int b = 0; b < 16; b++) {
                if ((w[b >> 2] >> (8 * (b & 3))) & 0xffu) {
                    const int64_t j = jl + b;
                    C[pos] = (int32_t)j;
                    if (V) V[pos] = acc[i < j ? packed_at(N, i, j) : packed_at(N, j, i)];
                    pos++;
                }
            }
        }
        base += total;
    }
}

// ------------------------------------------------------------ exact averages
// Fisher z of one float r in the reference's arithmetic (CorMat.cpp:36-42):
// r == 1 moves to the float below 1 - 1e-6, the quotient and its log are float,
// the half is taken in double and the result stored as float.
__device__ __forceinline__ float group_fisher(float r) {
    if (r == 1.0f) r = (float)((double)r - 1e-6);
    return (float)(0.5 * (double)logf((1.0f + r) / (1.0f - r)));
}

// Wave-aggregated append of this lane's pair when `take`.
__device__ __forceinline__ void append_pair(bool take, int i, int j, int2 *pairs, unsigned long long *count,
                                            int64_t cap, int lane) {
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int leader = __builtin_ctzll(m);
    unsigned long long b = 0;
    if (lane == leader) b = atomicAdd(count, (unsigned long long)__popcll(m));
    b = __shfl(b, leader, 64);
    if (take) {
        const unsigned long long k = b + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
        if ((int64_t)k < cap) pairs[k] = make_int2(i, j);
    }
}

// One wave per packed row i: pairs (i, j > i) whose average is within w of t or
// whose flag bit is set.  The row is read twice: the first pass counts its
// matches, so one atomic reserves the row's slots, and the second writes them
// in column order (a returning atomic per 64 columns had made the scan
// latency-bound).
__device__ __forceinline__ bool cand_take(const float *row, int64_t p0, int64_t q, int64_t n, float t, float w,
                                          const uint32_t *flags) {
    if (q >= n) return false;
    bool take = fabsf(row[q] - t) <= w;
    if (flags) {
        const int64_t p = p0 + q;
        take = take || ((flags[p >> 5] >> (p & 31)) & 1u);
    }
    return take;
}
__global__ __launch_bounds__(256) void group_cands_kernel(const float *__restrict__ acc, int N, float t, float w,
                                                          const uint32_t *__restrict__ flags,
                                                          int2 *__restrict__ pairs,
                                                          unsigned long long *__restrict__ count, int64_t cap)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N - 1) return;
    const int64_t p0 = packed_at(N, i, i + 1);
    const float *row = acc + p0;
    const int64_t n = N - 1 - i;
    unsigned long long c = 0;
    for (int64_t j0 = 0; j0 < n; j0 += 64) c += __popcll(__ballot(cand_take(row, p0, j0 + lane, n, t, w, flags)));
    if (c == 0) return;
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(count, c);
    b = __shfl(b, 0, 64);
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        const int64_t q = j0 + lane;
        const bool take = cand_take(row, p0, q, n, t, w, flags);
        const uint64_t m = __ballot(take);
        if (take) {
            const unsigned long long k = b + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
            if ((int64_t)k < cap) pairs[k] = make_int2((int)i, (int)(i + 1 + q));
        }
        b += (unsigned long long)__popcll(m);
    }
}

// Four pairs per wave, one per 16-lane group.  For each subject, r64 = sum_l
// z_i[l] z_j[l] with z recomputed in double from the series as the
// normalisation writes it, in the FP64 oracle's order: 64 partial sums, sum q
// taking l = q, q + 64, ... with fused multiply-add, then a halving tree (sum q
// adds sum q + s for s = 32, 16, ..., 1).  Lane g of a group holds sums g,
// g + 16, g + 32 and g + 48, so the tree's first two steps stay in the lane and
// the last four are shuffles inside the group.  Then the reference's float
// arithmetic: r as float, added (or its Fisher z) in subject order when it lies
// in [-1, 1], divided by S, mapped back for Fisher sums.
__global__ __launch_bounds__(256) void group_exact_kernel(const int2 *__restrict__ pairs,
                                                          const unsigned long long *__restrict__ count,
                                                          int64_t cap, const GroupSubj *__restrict__ subjs, int S,
                                                          int fisher, float *__restrict__ acc, int N)
{
    const int64_t n = (int64_t)min((unsigned long long)cap, *count);
    const int lane = threadIdx.x & 63, gl = lane & 15;
    const int64_t grp = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    const int64_t ngrp = ((int64_t)gridDim.x * blockDim.x) >> 4;
    // every group of a wave runs the same number of trips, so the shuffles
    // below always see all 64 lanes
    const int64_t trips = (n + ngrp - 1) / ngrp;
    for (int64_t t = 0; t < trips; t++) {
        const int64_t c = grp + t * ngrp;
        const bool live = c < n;
        const int2 pr = live ? pairs[c] : make_int2(0, 0);
        float sum = 0.0f;
        for (int s = 0; s < S; s++) {
            const GroupSubj g = subjs[s];
            const float *xi = g.xs + (int64_t)pr.x * g.ld;
            const float *xj = g.xs + (int64_t)pr.y * g.ld;
            const double mi = g.mu[pr.x], si = g.sd[pr.x], mj = g.mu[pr.y], sj = g.sd[pr.y];
            double p[4] = {0.0, 0.0, 0.0, 0.0};
            if (live) {
                for (int b = 0; b < g.L; b += 64) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int l = b + gl + 16 * k;
                        if (l < g.L) {
                            const double zi = ((double)xi[l] - mi) / si;
                            const double zj = ((double)xj[l] - mj) / sj;
                            p[k] = __builtin_fma(zi, zj, p[k]);
                        }
                    }
                }
            }
            p[0] = p[0] + p[2];  // s = 32
            p[1] = p[1] + p[3];
            double q = p[0] + p[1];  // s = 16
#pragma unroll
            for (int d = 8; d >= 1; d >>= 1) {
                const double o = __shfl_down(q, d, 16);
                if (gl < d) q = q + o;
            }
            const float r = (float)q;
            if (r >= -1.0f && r <= 1.0f) sum += fisher ? group_fisher(r) : r;
        }
        if (live && gl == 0) {
            float v = sum / (float)S;
            if (fisher) {
                const float e = expf(2.0f * v);
                v = (e - 1.0f) / (e + 1.0f);
            }
            acc[packed_at(N, pr.x, pr.y)] = v;
        }
    }
}

// One wave per row: the row's entries with a larger column.
__global__ __launch_bounds__(256) void csr_upper_pairs_kernel(const int32_t *__restrict__ R,
                                                              const int32_t *__restrict__ C, int N,
                                                              int2 *__restrict__ pairs,
                                                              unsigned long long *__restrict__ count, int64_t cap)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N) return;
    for (int64_t e0 = R[i]; e0 < R[i + 1]; e0 += 64) {
        const int64_t e = e0 + lane;
        const int j = e < R[i + 1] ? C[e] : -1;
        append_pair(j > (int)i, (int)i, j, pairs, count, cap, lane);
    }
}

__global__ __launch_bounds__(256) void group_refill_kernel(const int32_t *__restrict__ R,
                                                           const int32_t *__restrict__ C,
                                                           const float *__restrict__ acc, int N,
                                                           float *__restrict__ V)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= N) return;
    for (int64_t e = R[i] + lane; e < R[i + 1]; e += 64) {
        const int64_t j = C[e];
        V[e] = acc[i < j ? packed_at(N, i, j) : packed_at(N, j, i)];
    }
}

static unsigned stream_blocks(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

hipError_t launch_group_finish(float *acc, int64_t M1, int S, int fisher, hipStream_t st)
{
    if (M1 <= 0) return hipSuccess;
    if (S < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_finish_kernel, dim3(stream_blocks(M1)), dim3(256), 0, st, acc, M1, (float)S, fisher);
    return hipGetLastError();
}

hipError_t launch_group_digits(const float *acc, int64_t M1, uint32_t prefix, uint32_t mask, int shift,
                               unsigned long long *hist, hipStream_t st)
{
    if (M1 <= 0) return hipSuccess;
    if (shift < 0 || shift > 24) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_digits_kernel, dim3(stream_blocks(M1)), dim3(256), 0, st, acc, M1, prefix, mask, shift,
                       hist);
    return hipGetLastError();
}

hipError_t launch_group_ties(const float *acc, int N, float val, int64_t *tie_rows, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    hipLaunchKernelGGL(group_ties_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, acc, N, val, tie_rows);
    return hipGetLastError();
}

hipError_t launch_group_keep(const float *acc, int N, int64_t ld, float thr, int dens, int64_t p_cut, uint8_t *keep,
                             unsigned int *flag, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    const unsigned nbk = (unsigned)((N + 63) / 64);
    if (nbk > 65535 || ld % 64 || ld < (int64_t)nbk * 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_keep_kernel, dim3(nbk, nbk), dim3(256), 0, st, acc, N, ld, thr, dens, p_cut, keep, flag);
    return hipGetLastError();
}

hipError_t launch_group_rowcount(const uint8_t *keep, int N, int64_t ld, int64_t *cnt, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    if (ld % 64 || ld < N) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_rowcount_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, keep, N, ld, cnt);
    return hipGetLastError();
}

hipError_t launch_group_fill(const uint8_t *keep, int64_t ld, const float *acc, int N, const int32_t *R, int32_t *C,
                             float *V, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    if (ld % 64 || ld < N) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_fill_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, keep, ld, acc, N, R, C,
                       V);
    return hipGetLastError();
}

hipError_t launch_group_cands(const float *acc, int N, float t, float w, const uint32_t *flags, int2 *pairs,
                              unsigned long long *count, int64_t cap, hipStream_t st)
{
    if (N < 2) return hipSuccess;
    hipLaunchKernelGGL(group_cands_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, acc, N, t, w, flags,
                       pairs, count, cap);
    return hipGetLastError();
}

hipError_t launch_group_exact(const int2 *pairs, const unsigned long long *count, int64_t cap,
                              const GroupSubj *subjs, int S, int fisher, float *acc, int N, hipStream_t st)
{
    if (S < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(group_exact_kernel, dim3(4096), dim3(256), 0, st, pairs, count, cap, subjs, S, fisher, acc, N);
    return hipGetLastError();
}

hipError_t launch_csr_upper_pairs(const int32_t *R, const int32_t *C, int N, int2 *pairs,
                                  unsigned long long *count, int64_t cap, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    hipLaunchKernelGGL(csr_upper_pairs_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, R, C, N, pairs,
                       count, cap);
    return hipGetLastError();
}

hipError_t launch_group_refill(const int32_t *R, const int32_t *C, const float *acc, int N, float *V, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    hipLaunchKernelGGL(group_refill_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, st, R, C, acc, N, V);
    return hipGetLastError();
}

}  // namespace bna
