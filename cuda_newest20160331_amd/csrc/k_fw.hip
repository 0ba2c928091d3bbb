// k_fw.hip -- all-pairs shortest paths of a weighted graph by blocked
// Floyd-Warshall on a dense cost matrix held in HBM, for the weighted
// characteristic path length and the nodal efficiency of the CUBFW_Lp tool
// (W/src/Lp/CUBFW_Lp/CUBFW_Lp/CUBFW_Lp.cpp:13-162).
//
// The reference forms the cost matrix on the host (FormBlock.cpp:45-75: +inf
// everywhere, 1 / w for every CSR entry, 0 on the diagonal), runs a blocked
// Floyd-Warshall over it with blocks of 3072 streamed through the GPU, and sums
// 1.0f / d over the ordered pairs.  Here the whole matrix (Np x Np floats, Np =
// N rounded up to FW_B) stays in device memory and every block step is three
// launches:
//   1. the diagonal block (K, K) runs the plain recurrence over its own k;
//   2. the row panel (K, J) and the column panel (I, K) blocks run it with the
//      finished diagonal block on the other side;
//   3. every other block (I, J) takes min(D(I, J), D(I, K) (min,+) D(K, J)),
//      a min-plus product on the vector ALU with both panels staged in LDS.
// Each update is min(D(i, j), fl(D(i, k) + D(k, j))).  Costs are non-negative,
// so row and column k do not change at step k, and within a step the order of
// the updates does not matter; a min is exact, so the order of the k terms in
// step 3 does not matter either.  The float distances are therefore fixed by
// the block size FW_B alone (which k values a phase takes from a finished
// block), not by the launch shapes, and a CPU restatement of the same blocked
// recurrence reproduces them bit for bit.
#include "common.hpp"

#include <algorithm>

namespace bna {

constexpr int TP = FW_B + 4;  // LDS row pitch of a staged block, floats (rows stay 16-byte aligned)

// D = +inf off the diagonal, 0 on it (the diagonal of the padding included).
__global__ __launch_bounds__(256) void fw_fill_kernel(float *__restrict__ D, int Np)
{
    const int64_t n4 = (int64_t)Np * Np / 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const float inf = __builtin_inff();
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
        const int64_t e = q * 4;
        const int64_t i = e / Np, j = e - i * Np;
        float4 v = {inf, inf, inf, inf};
        if (i >= j && i < j + 4) {
            if (i == j) v.x = 0.f;
            else if (i == j + 1) v.y = 0.f;
            else if (i == j + 2) v.z = 0.f;
            else v.w = 0.f;
        }
        reinterpret_cast<float4 *>(D)[q] = v;
    }
}

// One thread per row, entries in CSR order (a repeated column keeps its last
// value, as FormBlock.cpp:52-66 does): D(i, j) = 1.0f / w, or 1 for a graph
// without weights.  Self loops are left at 0 (the reference zeroes the diagonal
// after the scatter).  bad |= 1 for a weight with the sign bit set or a NaN,
// |= 2 for a column outside [0, N).
__global__ __launch_bounds__(256) void fw_scatter_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                         const float *__restrict__ V, int N, int Np,
                                                         float *__restrict__ D, unsigned *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    unsigned flag = 0;
    float *row = D + (int64_t)i * Np;
    for (int e = R[i]; e < R[i + 1]; e++) {
        const int j = C[e];
        if (j < 0 || j >= N) {
            flag |= 2u;
            continue;
        }
        float c = 1.f;
        if (V) {
            const float w = V[e];
            if (__builtin_signbit(w) || w != w) flag |= 1u;
            c = 1.0f / w;
        }
        if (j != i) row[j] = c;
    }
    if (flag) atomicOr(bad, flag);
}

// The 8 x 8 block of a 128 x 128 tile that thread t of 256 owns: rows
// (t >> 4) * 8 .., columns (t & 15) * 8 ...
struct Own {
    float v[8][8];
};

__device__ __forceinline__ void load_own(Own &o, const float *__restrict__ D, int Np, int64_t r0, int64_t c0,
                                         int i0, int j0)
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const float4 *p = reinterpret_cast<const float4 *>(D + (r0 + i0 + r) * Np + c0 + j0);
        const float4 a = p[0], b = p[1];
        o.v[r][0] = a.x, o.v[r][1] = a.y, o.v[r][2] = a.z, o.v[r][3] = a.w;
        o.v[r][4] = b.x, o.v[r][5] = b.y, o.v[r][6] = b.z, o.v[r][7] = b.w;
    }
}

__device__ __forceinline__ void store_own(const Own &o, float *__restrict__ D, int Np, int64_t r0, int64_t c0,
                                          int i0, int j0)
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        float4 *p = reinterpret_cast<float4 *>(D + (r0 + i0 + r) * Np + c0 + j0);
        p[0] = make_float4(o.v[r][0], o.v[r][1], o.v[r][2], o.v[r][3]);
        p[1] = make_float4(o.v[r][4], o.v[r][5], o.v[r][6], o.v[r][7]);
    }
}

__device__ __forceinline__ void own_to_lds(const Own &o, float (*T)[TP], int i0, int j0)
{
#pragma unroll
    for (int r = 0; r < 8; r++) {
        float4 *p = reinterpret_cast<float4 *>(&T[i0 + r][j0]);
        p[0] = make_float4(o.v[r][0], o.v[r][1], o.v[r][2], o.v[r][3]);
        p[1] = make_float4(o.v[r][4], o.v[r][5], o.v[r][6], o.v[r][7]);
    }
}

// The sequential recurrence over the 128 k of block K inside one tile held as
// 8 x 8 register blocks.  MODE 0: the diagonal block, both row k and column k
// come from the tile itself; MODE 1: a row panel (K, J), column k from the
// finished diagonal block G; MODE 2: a column panel (I, K), row k from G.  The
// tile's own row / column k + 1 is published to T by its owners at the end of
// step k; one barrier per step (step k reads only row k and column k, which
// step k does not change, and the cells of row k + 1 / column k + 1 that step
// k reads hold the same value before and after it).
template <int MODE>
__device__ __forceinline__ void fw_steps(Own &o, float (*T)[TP], const float (*G)[TP], int bi, int bj)
{
    const int i0 = bi * 8, j0 = bj * 8;
    for (int kb = 0; kb < FW_B; kb += 8) {
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            const int k = kb + kk;
            float col[8], row[8];
#pragma unroll
            for (int r = 0; r < 8; r++) col[r] = MODE == 1 ? G[i0 + r][k] : T[i0 + r][k];
            {
                const float(*src)[TP] = MODE == 2 ? G : (const float(*)[TP])T;
                const float4 a = *reinterpret_cast<const float4 *>(&src[k][j0]);
                const float4 b = *reinterpret_cast<const float4 *>(&src[k][j0 + 4]);
                row[0] = a.x, row[1] = a.y, row[2] = a.z, row[3] = a.w;
                row[4] = b.x, row[5] = b.y, row[6] = b.z, row[7] = b.w;
            }
#pragma unroll
            for (int r = 0; r < 8; r++)
#pragma unroll
                for (int c = 0; c < 8; c++) o.v[r][c] = fminf(o.v[r][c], col[r] + row[c]);
            // publish row k + 1 (MODE 0, 1) and column k + 1 (MODE 0, 2)
            const int kn = k + 1;
            if (kn < FW_B) {
                if (MODE != 2 && (kn >> 3) == bi) {
                    const int rr = (kk + 1) & 7;
                    float4 *p = reinterpret_cast<float4 *>(&T[kn][j0]);
                    p[0] = make_float4(o.v[rr][0], o.v[rr][1], o.v[rr][2], o.v[rr][3]);
                    p[1] = make_float4(o.v[rr][4], o.v[rr][5], o.v[rr][6], o.v[rr][7]);
                }
                if (MODE != 1 && (kn >> 3) == bj) {
                    const int cc = (kk + 1) & 7;
#pragma unroll
                    for (int r = 0; r < 8; r++) T[i0 + r][kn] = o.v[r][cc];
                }
            }
            __syncthreads();
        }
    }
}

// Phase 1: the diagonal block (K, K).
__global__ __launch_bounds__(256) void fw_diag_kernel(float *__restrict__ D, int Np, int K)
{
    __shared__ __attribute__((aligned(16))) float T[FW_B][TP];
    const int t = threadIdx.x, bi = t >> 4, bj = t & 15;
    const int64_t o0 = (int64_t)K * FW_B;
    Own o;
    load_own(o, D, Np, o0, o0, bi * 8, bj * 8);
    own_to_lds(o, T, bi * 8, bj * 8);
    __syncthreads();
    fw_steps<0>(o, T, T, bi, bj);
    store_own(o, D, Np, o0, o0, bi * 8, bj * 8);
}

// Phase 2: workgroup w < nb - 1 is the row panel (K, J), the others the column
// panels (I, K); J, I run over the blocks other than K.
__global__ __launch_bounds__(256) void fw_panel_kernel(float *__restrict__ D, int Np, int K, int nb)
{
    __shared__ __attribute__((aligned(16))) float T[FW_B][TP];
    __shared__ __attribute__((aligned(16))) float G[FW_B][TP];
    const int t = threadIdx.x, bi = t >> 4, bj = t & 15;
    const bool rowp = (int)blockIdx.x < nb - 1;
    int X = rowp ? (int)blockIdx.x : (int)blockIdx.x - (nb - 1);
    if (X >= K) X++;
    const int64_t k0 = (int64_t)K * FW_B, x0 = (int64_t)X * FW_B;
    const int64_t r0 = rowp ? k0 : x0, c0 = rowp ? x0 : k0;
    // the finished diagonal block
    for (int q = t; q < FW_B * FW_B / 4; q += 256) {
        const int r = q / (FW_B / 4), c = (q % (FW_B / 4)) * 4;
        *reinterpret_cast<float4 *>(&G[r][c]) = *reinterpret_cast<const float4 *>(D + (k0 + r) * Np + k0 + c);
    }
    Own o;
    load_own(o, D, Np, r0, c0, bi * 8, bj * 8);
    own_to_lds(o, T, bi * 8, bj * 8);
    __syncthreads();
    if (rowp) fw_steps<1>(o, T, G, bi, bj);
    else fw_steps<2>(o, T, G, bi, bj);
    store_own(o, D, Np, r0, c0, bi * 8, bj * 8);
}

// Phase 3: block (I, J), I != K != J: D(I, J) = min(D(I, J), min over k in K of
// D(I, k) + D(k, J)).  P = D(I, K) and Q = D(K, J) pass through LDS in chunks
// of 32 k (35 KB), so three workgroups share a CU and the others compute while
// one loads.  Each thread keeps its 8 x 8 outputs in registers and reads, per
// 2 k, one float2 of P per row and two float4 of Q per k (few enough operand
// registers for three workgroups without spills).  P's float4 columns are
// XOR-swizzled by row block ((row >> 3) & 3): the four row blocks a wave reads
// at once then sit in different LDS banks.
constexpr int KC = 32;      // k per LDS chunk
constexpr int PP = KC + 4;  // P row pitch, floats

__device__ __forceinline__ int p_col4(int row, int c4) { return c4 ^ ((row >> 3) & 3); }

__device__ __forceinline__ void minplus_chunk(Own &o, const float (*P)[PP], const float (*Q)[FW_B], int i0, int j0)
{
#pragma unroll 2
    for (int k = 0; k < KC; k += 2) {
        float2 a[8];
#pragma unroll
        for (int r = 0; r < 8; r++)
            a[r] = *reinterpret_cast<const float2 *>(&P[i0 + r][p_col4(i0 + r, k >> 2) * 4 + (k & 3)]);
        float bq[2][8];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float4 x = *reinterpret_cast<const float4 *>(&Q[k + q][j0]);
            const float4 y = *reinterpret_cast<const float4 *>(&Q[k + q][j0 + 4]);
            bq[q][0] = x.x, bq[q][1] = x.y, bq[q][2] = x.z, bq[q][3] = x.w;
            bq[q][4] = y.x, bq[q][5] = y.y, bq[q][6] = y.z, bq[q][7] = y.w;
        }
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int c = 0; c < 8; c++) o.v[r][c] = fminf(fminf(o.v[r][c], a[r].x + bq[0][c]), a[r].y + bq[1][c]);
    }
}

// A thread's share of one chunk: 128 x 32 of P and 32 x 128 of Q are 1024
// float4 each, 4 per thread.
constexpr int SG4 = FW_B * KC / 4 / 256;

__device__ __forceinline__ void chunk_load(float (*P)[PP], float (*Q)[FW_B], const float *__restrict__ D, int Np,
                                           int64_t ri, int64_t kc, int64_t cj, int t)
{
    float4 sp[SG4], sq[SG4];
#pragma unroll
    for (int u = 0; u < SG4; u++) {
        const int q = t + 256 * u;
        sp[u] = *reinterpret_cast<const float4 *>(D + (ri + q / (KC / 4)) * Np + kc + (q % (KC / 4)) * 4);
        sq[u] = *reinterpret_cast<const float4 *>(D + (kc + (q >> 5)) * Np + cj + (q & 31) * 4);
    }
#pragma unroll
    for (int u = 0; u < SG4; u++) {
        const int q = t + 256 * u, r = q / (KC / 4);
        *reinterpret_cast<float4 *>(&P[r][p_col4(r, q % (KC / 4)) * 4]) = sp[u];
        *reinterpret_cast<float4 *>(&Q[q >> 5][(q & 31) * 4]) = sq[u];
    }
}

__global__ __launch_bounds__(256, 3) void fw_minplus_kernel(float *__restrict__ D, int Np, int K, int nb)
{
    __shared__ __attribute__((aligned(16))) float P[FW_B][PP];
    __shared__ __attribute__((aligned(16))) float Q[KC][FW_B];
    const int t = threadIdx.x, bi = t >> 4, bj = t & 15;
    const int i0 = bi * 8, j0 = bj * 8;
    // XCD-aware order: consecutive workgroups go to different XCDs, so XCD x
    // (= b mod 8) gets the contiguous run of tiles [x q + min(x, rem), ...) of
    // length q + (x < rem) in row-major order, and the workgroups it runs at
    // once share their P block in its L2.  A bijection of [0, m^2).
    const int m = nb - 1;
    const int64_t ntile = (int64_t)m * m, q8 = ntile / 8, rem = ntile % 8;
    const int64_t b = blockIdx.x, x = b & 7, l = b >> 3;
    const int64_t tile = x * q8 + min(x, rem) + l;
    int I = (int)(tile / m), J = (int)(tile % m);
    if (I >= K) I++;
    if (J >= K) J++;
    const int64_t k0 = (int64_t)K * FW_B, ri = (int64_t)I * FW_B, cj = (int64_t)J * FW_B;
    Own o;
    load_own(o, D, Np, ri, cj, i0, j0);
    for (int ch = 0; ch < FW_B / KC; ch++) {
        if (ch) __syncthreads();
        chunk_load(P, Q, D, Np, ri, k0 + ch * KC, cj, t);
        __syncthreads();
        minplus_chunk(o, P, Q, i0, j0);
    }
    store_own(o, D, Np, ri, cj, i0, j0);
}

// Per-row sums over the finished distances, one wave per 64 rows: for row u < N
// and v = 0 .. N - 1 in order, v != u, d = D(u, v) not below 0.1, the float sum
// li += 1.0f / d and the double sum rs += (double)(1.0f / d)
// (CUBFW_Lp.cpp:70-95); eff[u] = li / (N - 1), set to 0 below 1.0f / N
// (:97-109).  The 64 x 64 block of a step is read coalesced into LDS, and each
// lane then walks its own row of it.
__global__ __launch_bounds__(64) void fw_rows_kernel(const float *__restrict__ D, int Np, int N,
                                                     float *__restrict__ eff, double *__restrict__ rowsum)
{
    __shared__ __attribute__((aligned(16))) float S[64][68];
    const int lane = threadIdx.x;
    const int u0 = blockIdx.x * 64, u = u0 + lane;
    float li = 0.f;
    double rs = 0.0;
    for (int v0 = 0; v0 < N; v0 += 64) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int r = q * 4 + (lane >> 4), c = (lane & 15) * 4;
            // rows past N read padding rows of D (inside the Np x Np matrix)
            *reinterpret_cast<float4 *>(&S[r][c]) =
                *reinterpret_cast<const float4 *>(D + (int64_t)(u0 + r) * Np + v0 + c);
        }
        __syncthreads();
        const int nv = min(64, N - v0);
        for (int c = 0; c < nv; c++) {
            const float d = S[lane][c];
            if (v0 + c != u && !((double)d < 0.1)) {
                const float w = 1.0f / d;
                li += w;
                rs += (double)w;
            }
        }
        __syncthreads();
    }
    if (u < N) {
        float e = li / (float)(N - 1);
        if (e < 1.0f / (float)N) e = 0.f;
        eff[u] = e;
        rowsum[u] = rs;
    }
}

hipError_t launch_fw_init(const int32_t *R, const int32_t *C, const float *V, int N, int Np, float *D, unsigned *bad,
                          hipStream_t st)
{
    const int64_t n4 = (int64_t)Np * Np / 4;
    const int grid = (int)std::min<int64_t>((n4 + 255) / 256, 256 * 64);
    fw_fill_kernel<<<grid, 256, 0, st>>>(D, Np);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    fw_scatter_kernel<<<(N + 255) / 256, 256, 0, st>>>(R, C, V, N, Np, D, bad);
    return hipGetLastError();
}

hipError_t launch_fw_block(float *D, int Np, int K, hipStream_t st)
{
    const int nb = Np / FW_B;
    fw_diag_kernel<<<1, 256, 0, st>>>(D, Np, K);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nb == 1) return e;
    fw_panel_kernel<<<2 * (nb - 1), 256, 0, st>>>(D, Np, K, nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    fw_minplus_kernel<<<(nb - 1) * (nb - 1), 256, 0, st>>>(D, Np, K, nb);
    return hipGetLastError();
}

hipError_t launch_fw_rows(const float *D, int Np, int N, float *eff, double *rowsum, hipStream_t st)
{
    fw_rows_kernel<<<(N + 63) / 64, 64, 0, st>>>(D, Np, N, eff, rowsum);
    return hipGetLastError();
}

}  // namespace bna
