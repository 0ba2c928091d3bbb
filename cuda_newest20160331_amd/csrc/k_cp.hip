// k_cp.hip -- weighted clustering coefficient over a device CSR, the CUCP tool's
// Cp_2 (W/Cp_GPU/CUCP/Cp.cu:111-471).
//
// The reference's steps, restated:
//   inv_i  = 0 if t < 2 else (float)(1.0 / (t * (t - 1))), with the float
//            t = (float)R[i+1] - (float)R[i]                      (cal_k_kernel, :15-31)
//   alpha  = (float)(1.0 / V[m]), m the first entry of largest |V| (cublasIsamax, :189-205)
//   w      = (float)pow((double)(V * alpha), 1.0 / 3)              (cublasSscal + pow3_vector, :33-46, :206-213)
//   cci_i  = (W^3)_ii = sum_j w_ij sum_k w_jk w_ki                  (csrmm over dense row blocks + calc_cci_block, :48-92, :317-341)
//   Cp_i   = cci_i * inv_i                                           (dot_vv, :96-109, :451)
// The float t reproduces the reference's degree term bit for bit: past 2^24
// entries R rounds to a float and t can be off the row length by one or two.
//
// The reference builds dense blocks of 3072 rows of W and multiplies them with
// cuSPARSE.  Here one workgroup takes one row i and never forms a dense row: it
// marks row i's columns in an LDS bitmap (N bits) beside a per-word prefix count
// of set bits, so for every neighbour j and every k in row j one LDS read says
// whether (i, k) is an edge, and the prefix plus a popcount gives k's position in
// row i (columns ascend within a row) and so w_ik.  The LDS use is 8 bytes per
// 32 voxels (11 KB at N = 45000), so several workgroups share a CU.  Products
// are added in double in a fixed order (lane stride, then a fixed tree), so a
// row's cci is the same on every run.
#include "common.hpp"

#include <algorithm>

namespace bna {

constexpr int CP_THREADS = 256;

// Largest |V| with the smallest index on ties: key = |v| bits << 32 | ~index.
// NaN entries take no part.
__global__ __launch_bounds__(256) void cp_amax_kernel(const float *__restrict__ V, int64_t nnz,
                                                      unsigned long long *__restrict__ key)
{
    unsigned long long best = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t ab = __builtin_bit_cast(uint32_t, V[q]) & 0x7fffffffu;
        if (ab > 0x7f800000u) continue;
        const unsigned long long k = ((unsigned long long)ab << 32) | (unsigned long long)(0xffffffffu - (uint32_t)q);
        best = k > best ? k : best;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(best, o);
        best = other > best ? other : best;
    }
    if ((threadIdx.x & 63) == 0 && best) atomicMax(key, best);
}

// w = (float)pow((double)(V * alpha), 1/3) with alpha = (float)(1.0 / V[m]).
__global__ __launch_bounds__(256) void cp_weights_kernel(const float *__restrict__ V, int64_t nnz,
                                                         const unsigned long long *__restrict__ key,
                                                         float *__restrict__ w)
{
    const unsigned long long k = *key;
    float alpha = __builtin_nanf("");
    if (k) {
        const float vm = V[0xffffffffu - (uint32_t)(k & 0xffffffffu)];
        alpha = (float)(1.0 / (double)vm);
    }
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const float s = V[q] * alpha;
        w[q] = (float)pow((double)s, 1.0 / 3);
    }
}

// One workgroup per row i (columns outside [0, N) are skipped).  Dynamic LDS: bits[W] then pre[W], W = ceil(N / 32).
// WEIGHTED false: every w is 1 (a binary graph).
template <bool WEIGHTED>
__global__ __launch_bounds__(CP_THREADS) void cp_row_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                            const float *__restrict__ w, int N, int W,
                                                            float *__restrict__ cp)
{
    extern __shared__ uint32_t sm[];
    uint32_t *bits = sm;
    uint32_t *pre = sm + W;
    __shared__ uint32_t wave_tot[CP_THREADS / 64];
    __shared__ double red[CP_THREADS / 64];
    const int i = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int q = t; q < W; q += CP_THREADS) bits[q] = 0u;
    __syncthreads();
    const int b = R[i], e = R[i + 1];
    for (int q = b + t; q < e; q += CP_THREADS) {
        const int k = C[q];
        if ((unsigned)k < (unsigned)N) atomicOr(&bits[k >> 5], 1u << (k & 31));
    }
    __syncthreads();
    // prefix counts: thread t owns words [w0, w1)
    const int per = (W + CP_THREADS - 1) / CP_THREADS;
    const int w0 = min(W, t * per), w1 = min(W, w0 + per);
    uint32_t s = 0;
    for (int q = w0; q < w1; ++q) s += __popc(bits[q]);
    uint32_t inc = s;  // inclusive scan inside the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(inc, o);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wave_tot[wv] = inc;
    __syncthreads();
    uint32_t base = inc - s;
    for (int v = 0; v < wv; ++v) base += wave_tot[v];
    for (int q = w0; q < w1; ++q) {
        pre[q] = base;
        base += __popc(bits[q]);
    }
    __syncthreads();
    // sum over j in row i, k in row j with (i, k) an edge: w_ij w_jk w_ik.
    // Each wave takes every fourth j; the next j and its row bounds are loaded
    // while the current row is walked, and a lane keeps four of row j's
    // columns in flight (q, q + 64, q + 128, q + 192).
    double acc = 0.0;
    int jj = b + wv;
    int jn = jj < e ? C[jj] : -1;
    int bn = 0, en = 0;
    if ((unsigned)jn < (unsigned)N) {
        bn = R[jn];
        en = R[jn + 1];
    }
    while (jj < e) {
        const int j = jn, bj = bn, ej = en;
        const double wij = WEIGHTED ? (double)w[jj] : 1.0;
        const int jj2 = jj + CP_THREADS / 64;
        jn = jj2 < e ? C[jj2] : -1;
        bn = en = 0;
        if ((unsigned)jn < (unsigned)N) {
            bn = R[jn];
            en = R[jn + 1];
        }
        if ((unsigned)j < (unsigned)N) {
            for (int q = bj + lane; q < ej; q += 256) {
                int kk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) kk[u] = q + 64 * u < ej ? C[q + 64 * u] : -1;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = kk[u];
                    if ((unsigned)k >= (unsigned)N) continue;
                    const uint32_t word = bits[k >> 5];
                    const uint32_t m = 1u << (k & 31);
                    if (word & m) {
                        if (WEIGHTED) {
                            const int rank = (int)(pre[k >> 5] + __popc(word & (m - 1u)));
                            acc += wij * ((double)w[q + 64 * u] * (double)w[b + rank]);
                        } else {
                            acc += 1.0;
                        }
                    }
                }
            }
        }
        jj = jj2;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (t == 0) {
        double cci = red[0];
        for (int v = 1; v < CP_THREADS / 64; ++v) cci += red[v];
        float deg = (float)R[i + 1];
        deg -= (float)R[i];
        const float inv = deg < 2.0f ? 0.0f : (float)(1.0 / (double)(deg * (deg - 1.0f)));
        cp[i] = (float)cci * inv;
    }
}

hipError_t launch_cp(const int32_t *R, const int32_t *C, const float *V, int N, int64_t nnz, float *w,
                     unsigned long long *key, float *cp, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    if (N > CP_MAX_N) return hipErrorInvalidValue;
    const int W = (N + 31) / 32;
    const size_t lds = sizeof(uint32_t) * 2 * (size_t)W;
    if (lds > 65536) {  // past the default dynamic LDS limit (up to 128 KB here)
        hipError_t e = hipFuncSetAttribute((const void *)(V && nnz > 0 ? cp_row_kernel<true> : cp_row_kernel<false>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (V && nnz > 0) {
        hipError_t e = hipMemsetAsync(key, 0, sizeof(unsigned long long), st);
        if (e != hipSuccess) return e;
        const unsigned blocks = (unsigned)std::min<int64_t>(4096, (nnz + 255) / 256);
        hipLaunchKernelGGL(cp_amax_kernel, dim3(blocks), dim3(256), 0, st, V, nnz, key);
        hipLaunchKernelGGL(cp_weights_kernel, dim3(blocks), dim3(256), 0, st, V, nnz, key, w);
        hipLaunchKernelGGL(cp_row_kernel<true>, dim3((unsigned)N), dim3(CP_THREADS), lds, st, R, C, w, N, W, cp);
    } else {
        hipLaunchKernelGGL(cp_row_kernel<false>, dim3((unsigned)N), dim3(CP_THREADS), lds, st, R, C, w, N, W, cp);
    }
    return hipGetLastError();
}

}  // namespace bna
