// k_modularity.hip -- Newman leading-eigenvector modularity on a device CSR
// (the CUmodularity tool, B/Newman_Modularity_GPU/CUmodularity/modularity_GPU.cu
// and Modularity_CPU.cpp of the binary 20160115 tree; the weighted tree's copy
// of the tool does not compile).
//
// The partition is repeated bisection, driven from the host as in the
// reference (Partition_GPU, modularity_GPU.cu:408-528): round r takes the group
// G of nodes labelled r, builds G's subgraph on the device (rows of G in
// ascending order, the neighbours inside G with their original indices), and
// runs the power iteration of Lead_Vector (Modularity_CPU.cpp:209-266) on the
// generalised modularity matrix of G shifted by beta:
//   y_a = sum_{j in N_G(a)} x_j - (v_k / 2 / M) k_a - (sumBG_a - beta) x_a,
//   v_k = sum_a x_a k_a,  sumBG_a = k^G_a - k_a K_G / 2 / M,
// x the normalised previous iterate (the first one is the start vector, not
// normalised), until |v - v0| <= eps or |v + v0| <= eps or max_iter iterations.
//
// One iteration is one launch of nm_iter_kernel.  It keeps the iterate
// unnormalised (y) and the inverse norms in the state, so iteration t reads
// x = y_{t-1} / n_{t-1} with one multiply per row.  The error norms of
// iteration t - 1 are the sums of (x - v0)^2 and (x + v0)^2 that iteration t
// forms row by row from y_{t-1} and y_{t-2}, the reference's own definition;
// the stopping rule is therefore applied one launch later, and a launch that
// finds the rule fired leaves the state at iteration t - 1.  The last
// workgroup to finish reduces the per-workgroup partial sums in a fixed order
// and advances the state, so the host only looks at a flag once per batch.
#include "common.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace bna {

struct NmState {
    double inv1;    // 1 / n of iteration t - 1 (1 for the start vector)
    double inv2;    // 1 / n of iteration t - 2
    double vk;      // sum_a x_a k_a of the coming iteration's x
    double n_prev;  // n of iteration t - 1
    double lambda;  // n of the final iteration
    double err1, err2;
    double eps;
    int32_t iter;   // the coming iteration's index t
    int32_t max_iter;
    int32_t done;
    int32_t final_iter;  // index of the final iteration once done
    uint32_t arrive;
    int32_t pad;
};

constexpr int NM_THREADS = 256;
constexpr double NM_LAMBDA = 0.01;  // a division needs an eigenvalue above this (modularity_GPU.cu:30)
constexpr int NM_MIN_GROUP = 1;     // and more than this many nodes on each side (:31)
constexpr int NM_GPU_ROUND = 25000; // groups above this size take Sub_Partition_GPU's shift rule (:494)
constexpr int NM_ROWS_PER_BLOCK = NM_THREADS / 64;
static_assert(sizeof(NmState) <= 128, "two state slots fit the pinned stage");

__device__ inline double nm_wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// cnt[a] = neighbours of node AD[a] in its own group (the same label).  One
// wave per row.
__global__ __launch_bounds__(256) void nm_count_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                       const int32_t *__restrict__ AD, int n,
                                                       const int32_t *__restrict__ res, int32_t *__restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= n) return;
    const int i = AD[a];
    const int round = res[i];
    int c = 0;
    for (int k = R[i] + lane; k < R[i + 1]; k += 64) c += res[C[k]] == round;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) cnt[a] = c;
}

// SC[SR[a] ..] = the neighbours of AD[a] in its own group, in CSR order.
__global__ __launch_bounds__(256) void nm_scatter_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                         const int32_t *__restrict__ AD, int n,
                                                         const int32_t *__restrict__ res,
                                                         const int32_t *__restrict__ SR, int32_t *__restrict__ SC)
{
    const int lane = threadIdx.x & 63;
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= n) return;
    const int i = AD[a];
    const int round = res[i];
    int out = SR[a];
    for (int k0 = R[i]; k0 < R[i + 1]; k0 += 64) {
        const int k = k0 + lane;
        const int c = k < R[i + 1] ? C[k] : 0;
        const bool in = k < R[i + 1] && res[c] == round;
        const uint64_t m = __ballot(in);
        if (in) SC[out + __popcll(m & ((1ull << lane) - 1ull))] = c;
        out += __popcll(m);
    }
}

// sumBG (Sub_Partition, Modularity_CPU.cpp:179-183) of one group's rows.
__global__ __launch_bounds__(256) void nm_sumbg_range_kernel(const int32_t *__restrict__ R,
                                                             const int32_t *__restrict__ AD,
                                                             const int32_t *__restrict__ SR, int n, double innerM,
                                                             double m, double *__restrict__ sumBG)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int i = AD[a];
    sumBG[a] = (double)(SR[a + 1] - SR[a]) - (double)(R[i + 1] - R[i]) * innerM / 2 / m;
}

// The start vector, the node's index (W/Newman_Modularity_GPU/CUmodularity/
// modularity_GPU.cu:243), in the compact buffer y and the original-index
// buffer yg.
__global__ __launch_bounds__(256) void nm_start_kernel(const int32_t *__restrict__ AD, int n, double *__restrict__ y,
                                                       double *__restrict__ yg)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int i = AD[a];
    y[a] = (double)i;
    yg[i] = (double)i;
}

__global__ __launch_bounds__(NM_THREADS) void nm_iter_kernel(const int32_t *__restrict__ R,
                                                             const int32_t *__restrict__ SR,
                                                             const int32_t *__restrict__ SC,
                                                             const int32_t *__restrict__ AD,
                                                             const double *__restrict__ sumBG, int n, double m,
                                                             double beta, const double *__restrict__ yg_prev,
                                                             double *__restrict__ yg_cur,
                                                             const double *__restrict__ y_prev,
                                                             const double *__restrict__ y_prev2,
                                                             double *__restrict__ y_cur, double *__restrict__ part,
                                                             NmState *__restrict__ st)
{
    if (st->done) return;
    __shared__ double red[4][NM_ROWS_PER_BLOCK];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double inv1 = st->inv1, inv2 = st->inv2, vk = st->vk;
    const bool check = st->iter > 0;
    double syy = 0.0, syk = 0.0, e1 = 0.0, e2 = 0.0;
    for (int a = blockIdx.x * NM_ROWS_PER_BLOCK + w; a < n; a += gridDim.x * NM_ROWS_PER_BLOCK) {
        const int rb = SR[a], re = SR[a + 1];
        double s = 0.0;
        for (int t0 = rb + lane; t0 < re; t0 += 4 * 64) {
            // unconditional loads (a slot past the row end rereads its last
            // entry) so that the four are in flight together
            int cc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) cc[u] = SC[min(t0 + 64 * u, re - 1)];
            double xx[4];
#pragma unroll
            for (int u = 0; u < 4; u++) xx[u] = yg_prev[cc[u]];
#pragma unroll
            for (int u = 0; u < 4; u++) s += (t0 + 64 * u < re) ? xx[u] : 0.0;
        }
        s = nm_wave_sum(s);
        if (lane == 0) {
            const int i = AD[a];
            const double k = (double)(R[i + 1] - R[i]);
            const double x0 = y_prev[a] * inv1;
            const double y = s * inv1 - (vk / 2 / m * k + (sumBG[a] - beta) * x0);
            y_cur[a] = y;
            yg_cur[i] = y;
            syy += y * y;
            syk += y * k;
            if (check) {
                const double v0 = y_prev2[a] * inv2;
                e1 += (x0 - v0) * (x0 - v0);
                e2 += (x0 + v0) * (x0 + v0);
            }
        }
    }
    if (lane == 0) {
        red[0][w] = syy;
        red[1][w] = syk;
        red[2][w] = e1;
        red[3][w] = e2;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int k = 0; k < NM_ROWS_PER_BLOCK; k++) t += red[threadIdx.x][k];
        part[4 * blockIdx.x + threadIdx.x] = t;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) is_last = (atomicAdd(&st->arrive, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < (int)gridDim.x; b += NM_THREADS)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += part[4 * b + q];
#pragma unroll
    for (int q = 0; q < 4; q++) acc[q] = nm_wave_sum(acc[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) red[q][w] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < NM_ROWS_PER_BLOCK; k++)
#pragma unroll
            for (int q = 0; q < 4; q++) t[q] += red[q][k];
        st->arrive = 0;
        const int it = st->iter;
        if (it > 0) {
            // iteration it - 1 is complete: the reference's loop test
            st->err1 = sqrt(t[2]);
            st->err2 = sqrt(t[3]);
            if (!(st->err1 > st->eps && st->err2 > st->eps && it < st->max_iter)) {
                st->done = 1;
                st->final_iter = it - 1;
                st->lambda = st->n_prev;
                return;
            }
        }
        const double nt = sqrt(t[0]);
        st->inv2 = st->inv1;
        st->inv1 = 1.0 / nt;
        st->n_prev = nt;
        st->vk = t[1] / nt;
        st->iter = it + 1;
    }
}

// Entries (i, j) of the whole graph with res[i] == res[j], added to *same.
__global__ __launch_bounds__(256) void nm_same_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                      int N, const int32_t *__restrict__ res,
                                                      unsigned long long *__restrict__ same)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N) return;
    const int li = res[i];
    unsigned long long c = 0;
    for (int k = R[i] + lane; k < R[i + 1]; k += 64) c += res[C[k]] == li;
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0 && c) atomicAdd(same, c);
}

// ---- one workgroup per group: the whole of Sub_Partition in one launch
// Groups of at most NM_SMALL_ROWS nodes are partitioned by one workgroup each,
// all groups of a batch side by side: the workgroup runs the power iteration
// (one thread per row, the row's neighbours added in CSR order), the shifted
// second run when the first eigenvalue is negative, and the split counts, with
// only workgroup barriers between iterations.  Its arithmetic is that of
// nm_iter_kernel.
constexpr int NM_SMALL_ROWS = 8192;
constexpr int NM_GT = 1024;  // threads of the group kernel

__device__ inline double nm_block_sum(double v, double *sh)
{
    v = nm_wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < NM_GT / 64; k++) t += sh[k];  // every thread, the same fixed order
    return t;
}

__device__ inline long long nm_block_count(long long v, long long *sh)
{
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    long long t = 0;
    for (int k = 0; k < NM_GT / 64; k++) t += sh[k];
    return t;
}

__global__ __launch_bounds__(NM_GT) void nm_group_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ SR,
                                                         const int32_t *__restrict__ SC, const int32_t *__restrict__ AD,
                                                         const NmGroup *__restrict__ groups, double m, int max_iter,
                                                         double eps, double *__restrict__ sumBG, double *y0, double *y1,
                                                         double *y2, double *yg0, double *yg1,
                                                         int8_t *__restrict__ side, NmGroupOut *__restrict__ out)
{
    __shared__ double shd[NM_GT / 64];
    __shared__ long long shl[NM_GT / 64];
    __shared__ double shv[NM_GT / 64];
    __shared__ int shi[NM_GT / 64];
    const NmGroup gr = groups[blockIdx.x];
    const int a0 = gr.a0, n = gr.a1 - gr.a0;
    const int tid = threadIdx.x;
    double *y[3] = {y0 + a0, y1 + a0, y2 + a0};
    double *yg[2] = {yg0, yg1};
    const int32_t *ad = AD + a0;
    const int32_t *sr = SR + a0;
    double *sb = sumBG + a0;
    for (int a = tid; a < n; a += NM_GT) {
        const int i = ad[a];
        sb[a] = (double)(sr[a + 1] - sr[a]) - (double)(R[i + 1] - R[i]) * gr.inner / 2 / m;
    }
    int iters_all = 0;
    // one Lead_Vector run; returns the signed eigenvalue and leaves v in y[*fin]
    // scaled by *inv (v = y * inv)
    auto lead = [&](double beta, int &fin, double &inv) -> double {
        double vk = 0.0;
        for (int a = tid; a < n; a += NM_GT) {
            const int i = ad[a];
            y[2][a] = (double)i;
            yg[1][i] = (double)i;
            vk += (double)i * (double)(R[i + 1] - R[i]);
        }
        vk = nm_block_sum(vk, shd);  // its barriers also publish the start vector
        double inv1 = 1.0, inv2 = 1.0, n_prev = 0.0;
        int T = -1;
        for (int t = 0;; t++) {
            const double *ygp = yg[(t + 1) & 1];
            double *ygc = yg[t & 1];
            const double *yp = y[(t + 2) % 3], *yp2 = y[(t + 1) % 3];
            double *yc = y[t % 3];
            double syy = 0.0, syk = 0.0, e1 = 0.0, e2 = 0.0;
            for (int a = tid; a < n; a += NM_GT) {
                const int rb = sr[a], re = sr[a + 1];
                double sum = 0.0;
                int k = rb;
                for (; k + 4 <= re; k += 4) {
                    const double q0 = ygp[SC[k]], q1 = ygp[SC[k + 1]], q2 = ygp[SC[k + 2]], q3 = ygp[SC[k + 3]];
                    sum += q0;
                    sum += q1;
                    sum += q2;
                    sum += q3;
                }
                for (; k < re; k++) sum += ygp[SC[k]];
                const int i = ad[a];
                const double kd = (double)(R[i + 1] - R[i]);
                const double x0 = yp[a] * inv1;
                const double yy = sum * inv1 - (vk / 2 / m * kd + (sb[a] - beta) * x0);
                yc[a] = yy;
                ygc[i] = yy;
                syy += yy * yy;
                syk += yy * kd;
                if (t > 0) {
                    const double v0 = yp2[a] * inv2;
                    e1 += (x0 - v0) * (x0 - v0);
                    e2 += (x0 + v0) * (x0 + v0);
                }
            }
            const double Syy = nm_block_sum(syy, shd), Syk = nm_block_sum(syk, shd);
            const double E1 = nm_block_sum(e1, shd), E2 = nm_block_sum(e2, shd);
            if (t > 0) {
                const double err1 = sqrt(E1), err2 = sqrt(E2);
                if (!(err1 > eps && err2 > eps && t < max_iter)) {
                    T = t - 1;
                    break;
                }
            }
            const double nt = sqrt(Syy);
            inv2 = inv1;
            inv1 = 1.0 / nt;
            n_prev = nt;
            vk = Syk / nt;
        }
        iters_all += T + 1;
        fin = T % 3;
        inv = inv1;
        // sign: v . v0 at the first largest |v| (Modularity_CPU.cpp:268-272)
        const double *yT = y[T % 3], *yP = y[(T + 2) % 3];
        double best = -1.0;
        int bi = 0x7fffffff;
        for (int a = tid; a < n; a += NM_GT) {
            const double av = fabs(yT[a] * inv1);
            if (av > best) {
                best = av;
                bi = a;
            }
        }
        for (int d = 32; d >= 1; d >>= 1) {
            const double ob = __shfl_xor(best, d, 64);
            const int oi = __shfl_xor(bi, d, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        __syncthreads();
        if ((tid & 63) == 0) {
            shv[tid >> 6] = best;
            shi[tid >> 6] = bi;
        }
        __syncthreads();
        best = shv[0];
        bi = shi[0];
        for (int k = 1; k < NM_GT / 64; k++)
            if (shv[k] > best || (shv[k] == best && shi[k] < bi)) {
                best = shv[k];
                bi = shi[k];
            }
        const double vm = yT[bi] * inv1, v0m = yP[bi] * inv2;
        return (vm * v0m > 0) ? n_prev : -n_prev;
    };
    int fin = 0;
    double inv = 1.0;
    double lambda = lead(0.0, fin, inv);
    // a group of at most NM_SMALL_ROWS nodes takes Sub_Partition's rule
    if (lambda < 0.0) {
        __syncthreads();
        lambda += lead(-lambda, fin, inv);
    }
    long long sp = 0, sn = 0;
    for (int a = tid; a < n; a += NM_GT) {
        const double v = y[fin][a] * inv;
        sp += v > 0;
        sn += v <= 0;
    }
    sp = nm_block_count(sp, shl);
    sn = nm_block_count(sn, shl);
    for (int a = tid; a < n; a += NM_GT) {
        const double v = y[fin][a] * inv;
        side[a0 + a] = (int8_t)(v * ((double)(sp - sn) + 0.5) <= 0);
    }
    if (tid == 0) {
        NmGroupOut o;
        o.lambda = lambda;
        o.subP = sp;
        o.subN = sn;
        o.iters = iters_all;
        o.pad = 0;
        out[blockIdx.x] = o;
    }
}

// ---------------------------------------------------------------- host side
static int ceil_div(int64_t a, int b) { return (int)((a + b - 1) / b); }

size_t nm_state_bytes() { return sizeof(NmState); }

// One power iteration run of Lead_Vector on the group whose rows are
// [a0, a0 + n) of the batch (sumBG already set): returns the signed eigenvalue
// estimate, leaves v (normalised, host) in v_out.
static hipError_t nm_lead_vector(const NmBuffers &b, const int32_t *R, int a0, int n, double m, double beta,
                                 int max_iter, double eps, std::vector<double> &v_out, double &lambda, int64_t &iters,
                                 hipStream_t st)
{
    hipError_t e;
    const int32_t *AD = b.AD + a0, *SR = b.SR + a0;
    const double *sumBG = b.sumBG + a0;
    double *y[3] = {b.y[0] + a0, b.y[1] + a0, b.y[2] + a0};
    hipLaunchKernelGGL(nm_start_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, AD, n, y[2], b.yg[1]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // v_k of the start vector, in the reference's order (a sequential sum)
    NmState h{};
    {
        double vk = 0.0;
        for (int a = 0; a < n; a++) {
            const int i = b.AD_host[a0 + a];
            vk += (double)i * (double)(b.R_host[i + 1] - b.R_host[i]);
        }
        h.vk = vk;
    }
    h.inv1 = 1.0;
    h.inv2 = 1.0;
    h.eps = eps;
    h.max_iter = max_iter;
    if ((e = hipMemcpyAsync(b.state, &h, sizeof(NmState), hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    const int nblk = std::min(NM_BLOCKS_MAX, std::max(1, ceil_div(n, NM_ROWS_PER_BLOCK)));
    NmState *hs = (NmState *)b.state_host;
    // launches in batches, one state copy per batch; a batch is queued before
    // the host waits for the previous batch's copy
    const int batch = 32;
    int launched = 0;
    auto enqueue = [&](int slot) -> hipError_t {
        const int upto = std::min(launched + batch, max_iter + 1);
        for (int t = launched; t < upto; t++)
            hipLaunchKernelGGL(nm_iter_kernel, dim3(nblk), dim3(NM_THREADS), 0, st, R, SR, b.SC, AD, sumBG, n, m,
                               beta, b.yg[(t + 1) & 1], b.yg[t & 1], y[(t + 2) % 3], y[(t + 1) % 3], y[t % 3], b.part,
                               (NmState *)b.state);
        hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
        launched = upto;
        if ((r = hipMemcpyAsync(&hs[slot], b.state, sizeof(NmState), hipMemcpyDeviceToHost, st)) != hipSuccess)
            return r;
        return hipEventRecord(b.ev[slot], st);
    };
    // every pass t <= max_iter may be needed (pass t applies the test of
    // iteration t - 1), so at most max_iter + 1 launches
    int queued = 0, oldest_slot = 0;
    const NmState *fin = nullptr;
    if ((e = enqueue(0)) != hipSuccess) return e;
    queued = 1;
    while (queued > 0) {
        if (queued == 1 && launched < max_iter + 1) {
            if ((e = enqueue(oldest_slot ^ 1)) != hipSuccess) return e;
            queued = 2;
        }
        if ((e = hipEventSynchronize(b.ev[oldest_slot])) != hipSuccess) return e;
        queued--;
        if (hs[oldest_slot].done || queued == 0) {
            fin = &hs[oldest_slot];
            break;
        }
        oldest_slot ^= 1;
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (!fin || !fin->done) return hipErrorUnknown;  // the state machine always stops by max_iter + 1 passes
    const NmState f = *fin;
    const int T = f.final_iter;
    iters = T + 1;
    // v = y_T / n_T, v0 = y_{T-1} / n_{T-1}; the sign of the eigenvalue is that
    // of v0 . v at the largest |v| (Modularity_CPU.cpp:268-272)
    std::vector<double> yT(n), yP(n);
    if ((e = hipMemcpy(yT.data(), y[T % 3], sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    if ((e = hipMemcpy(yP.data(), y[(T + 2) % 3], sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess) return e;
    v_out.resize(n);
    const double invT = f.inv1, invP = f.inv2;
    int mi = 0;
    for (int a = 0; a < n; a++) {
        v_out[a] = yT[a] * invT;
        if (std::fabs(v_out[a]) > std::fabs(v_out[mi])) mi = a;
    }
    const double v0mi = yP[mi] * invP;
    lambda = (n > 0 && v_out[mi] * v0mi > 0) ? f.lambda : -f.lambda;
    return hipSuccess;
}

// The reference divides the group labelled Round, Round = 0, 1, 2, ... in
// turn; a division labels its halves Max + 1 and Max + 2.  A group's outcome
// depends on its members only, so every group that exists when a batch starts
// (labels lo .. hi) is divided side by side, and the outcomes are applied in
// label order afterwards, which numbers the halves exactly as the sequential
// loop does.  Groups above NM_SMALL_ROWS nodes take the multi-workgroup power
// iteration one after another; the others one workgroup each, in one launch.
hipError_t run_modularity(const int32_t *R, const int32_t *C, int N, int64_t nnz, int max_iter, double eps,
                          NmBuffers &b, int32_t *labels, double *Q, int *modules, int64_t *iters_total,
                          int *rounds_out, hipStream_t st)
{
    hipError_t e;
    const double m = (double)(nnz / 2);
    std::vector<int32_t> res(N, 0), adjust(2 * (size_t)N + 4, 0), cnt, SRh;
    int next = 0, max_result = 0, index = 1, mods = 1;
    int64_t its = 0;
    std::vector<double> v;
    std::vector<NmGroup> groups;
    std::vector<int> small;  // batch group index of each small group, in launch order
    std::vector<NmGroupOut> gout;
    std::vector<int8_t> side(N);
    std::vector<int> lab_of;  // label of each batch group
    while (next <= max_result) {
        const int lo = next, hi = max_result;
        // the batch's nodes grouped by label, ascending within a group
        std::vector<int> count(hi - lo + 1, 0);
        for (int i = 0; i < N; i++)
            if (res[i] >= lo) count[res[i] - lo]++;
        groups.clear();
        lab_of.clear();
        std::vector<int> start(hi - lo + 2, 0);
        for (int r = lo; r <= hi; r++) start[r - lo + 1] = start[r - lo] + count[r - lo];
        b.AD_host.assign(start[hi - lo + 1], 0);
        std::vector<int> fill(start.begin(), start.end() - 1);
        std::vector<double> inner(hi - lo + 1, 0.0);
        for (int i = 0; i < N; i++)
            if (res[i] >= lo) {
                b.AD_host[fill[res[i] - lo]++] = i;
                inner[res[i] - lo] += (double)(b.R_host[i + 1] - b.R_host[i]);
            }
        const int nb = (int)b.AD_host.size();
        for (int r = lo; r <= hi; r++)
            if (count[r - lo]) {
                groups.push_back(NmGroup{start[r - lo], start[r - lo + 1], inner[r - lo]});
                lab_of.push_back(r);
            }
        if ((e = hipMemcpyAsync(b.res, res.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st)) != hipSuccess)
            return e;
        if ((e = hipMemcpyAsync(b.AD, b.AD_host.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st)) !=
            hipSuccess)
            return e;
        hipLaunchKernelGGL(nm_count_kernel, dim3(ceil_div(nb, 4)), dim3(256), 0, st, R, C, b.AD, nb, b.res, b.cnt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        cnt.resize(nb);
        if ((e = hipMemcpyAsync(cnt.data(), b.cnt, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        SRh.resize(nb + 1);
        SRh[0] = 0;
        for (int a = 0; a < nb; a++) SRh[a + 1] = SRh[a] + cnt[a];
        if ((e = hipMemcpyAsync(b.SR, SRh.data(), sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice, st)) !=
            hipSuccess)
            return e;
        hipLaunchKernelGGL(nm_scatter_kernel, dim3(ceil_div(nb, 4)), dim3(256), 0, st, R, C, b.AD, nb, b.res, b.SR,
                           b.SC);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // per group: lambda, split counts and side flags
        const int ng = (int)groups.size();
        std::vector<double> lam(ng, 0.0);
        std::vector<long long> spv(ng, 0), snv(ng, 0);
        small.clear();
        std::vector<NmGroup> sg;
        for (int q = 0; q < ng; q++)
            if (groups[q].a1 - groups[q].a0 > NM_SMALL_ROWS) {
                const NmGroup &G = groups[q];
                const int n = G.a1 - G.a0;
                hipLaunchKernelGGL(nm_sumbg_range_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, R, b.AD + G.a0,
                                   b.SR + G.a0, n, G.inner, m, b.sumBG + G.a0);
                if ((e = hipGetLastError()) != hipSuccess) return e;
                double lambda = 0.0;
                int64_t it = 0;
                if ((e = nm_lead_vector(b, R, G.a0, n, m, 0.0, max_iter, eps, v, lambda, it, st)) != hipSuccess)
                    return e;
                its += it;
                // the shift rule of the path the reference takes for this group
                // size (Sub_Partition_GPU above 25000 nodes, Sub_Partition below)
                const bool shift = n > NM_GPU_ROUND ? lambda < -NM_LAMBDA : lambda < 0.0;
                if (shift) {
                    double l2 = 0.0;
                    if ((e = nm_lead_vector(b, R, G.a0, n, m, -lambda, max_iter, eps, v, l2, it, st)) != hipSuccess)
                        return e;
                    its += it;
                    lambda += l2;
                }
                long long sp = 0, sn = 0;
                for (int a = 0; a < n; a++) {
                    sp += v[a] > 0;
                    sn += v[a] <= 0;
                }
                for (int a = 0; a < n; a++) side[G.a0 + a] = (int8_t)(v[a] * ((double)(sp - sn) + 0.5) <= 0);
                lam[q] = lambda;
                spv[q] = sp;
                snv[q] = sn;
            } else {
                small.push_back(q);
                sg.push_back(groups[q]);
            }
        if (!sg.empty()) {
            const int ns = (int)sg.size();
            if ((e = hipMemcpyAsync(b.groups, sg.data(), sizeof(NmGroup) * ns, hipMemcpyHostToDevice, st)) !=
                hipSuccess)
                return e;
            hipLaunchKernelGGL(nm_group_kernel, dim3(ns), dim3(NM_GT), 0, st, R, b.SR, b.SC, b.AD, b.groups, m,
                               max_iter, eps, b.sumBG, b.y[0], b.y[1], b.y[2], b.yg[0], b.yg[1], b.side, b.gout);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            gout.resize(ns);
            if ((e = hipMemcpyAsync(gout.data(), b.gout, sizeof(NmGroupOut) * ns, hipMemcpyDeviceToHost, st)) !=
                hipSuccess)
                return e;
            std::vector<int8_t> sd(nb);
            if ((e = hipMemcpyAsync(sd.data(), b.side, (size_t)nb, hipMemcpyDeviceToHost, st)) != hipSuccess)
                return e;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
            for (int k = 0; k < ns; k++) {
                const int q = small[k];
                lam[q] = gout[k].lambda;
                spv[q] = gout[k].subP;
                snv[q] = gout[k].subN;
                its += gout[k].iters;
                for (int a = groups[q].a0; a < groups[q].a1; a++) side[a] = sd[a];
            }
        }
        // the outcomes in label order (Sub_Partition, Modularity_CPU.cpp:187-203)
        for (int q = 0; q < ng; q++) {
            const bool sub = lam[q] > NM_LAMBDA && spv[q] > NM_MIN_GROUP && snv[q] > NM_MIN_GROUP;
            if (sub) {
                for (int a = groups[q].a0; a < groups[q].a1; a++) res[b.AD_host[a]] = max_result + 1 + side[a];
                max_result += 2;
            } else {
                adjust[lab_of[q]] = index++;
            }
            mods += sub;
        }
        next = hi + 1;
    }
    // Q = (entries inside a module - sum over modules of (degree sum)^2 / 2M) / 2M
    if ((e = hipMemcpyAsync(b.res, res.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(b.same, 0, sizeof(unsigned long long), st)) != hipSuccess) return e;
    hipLaunchKernelGGL(nm_same_kernel, dim3(ceil_div(N, 4)), dim3(256), 0, st, R, C, N, b.res, b.same);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    unsigned long long same = 0;
    if ((e = hipMemcpyAsync(&same, b.same, sizeof(same), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    std::vector<double> ksum(2 * (size_t)N + 4, 0.0);
    for (int i = 0; i < N; i++) ksum[res[i]] += (double)(b.R_host[i + 1] - b.R_host[i]);
    double expect = 0.0;
    for (double k : ksum) expect += k * k / 2 / m;
    *Q = ((double)same - expect) / (2 * m);
    for (int i = 0; i < N; i++) labels[i] = adjust[res[i]];
    *modules = mods;
    if (iters_total) *iters_total = its;
    if (rounds_out) *rounds_out = max_result + 1;
    return hipSuccess;
}

}  // namespace bna
