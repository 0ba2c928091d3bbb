// k_paths.hip -- shortest-path measures of a binary graph on the device CSR:
// the characteristic path length (Lp) and betweenness centrality (BC).
//
// Lp follows CUBFS_Lp (B/src/Lp/CUBFS_Lp/CUBFS_Lp/CUBFS_LP.cpp:14-53,
// BFS_APSP_GPU.cu:7-50): the hop distance d(i, j) of every ordered pair from a
// BFS out of every vertex, then Lp = 1 / (sum over pairs with d > 0 of
// 1.0f / d, divided by N and by N - 1); unreachable pairs add nothing.  The
// reference stores the N x N distance matrix on the host.  Here only the number
// of ordered pairs at each distance is kept, from a bit-parallel multi-source
// BFS: 256 sources run together, each vertex holds one bit per source in four
// 64-bit words (visited, frontier, next), and one level of all 256 searches is
// one pass over the CSR: a vertex ORs its neighbours' frontier words and keeps
// the bits it has not seen.  The level's new bits, counted with popcount, are
// the pairs at that distance.
//
// BC follows CUBC's Brandes betweenness (B/src/BC/CUBC/CUBC.cpp:69-94 and the
// node/predecessor kernel it builds, BC_GPU_node_pred.cu): for every source s,
// sigma_s(v) = the number of shortest s-v paths, delta_s(v) = sum over the
// successors w of v (d(s, w) = d(s, v) + 1) of sigma_s(v) / sigma_s(w) *
// (1 + delta_s(w)), and BC(v) = sum over s != v of delta_s(v) (ordered pairs:
// an undirected graph's values are twice the unordered ones, as the reference's).
// One workgroup per source runs a level-synchronous BFS with a queue in BFS
// order; sigma is summed in double with atomics (integer counts below 2^53, so
// the order does not matter), delta is pulled from the successors in a fixed
// order, and the per-source deltas of a batch are added into BC in source
// order, so the result does not depend on scheduling.
#include "common.hpp"

#include <algorithm>

namespace bna {

typedef unsigned long long u64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- Lp
// Batch of sources [s0, s0 + ns), ns <= 256: clears the bit planes and sets
// source b's bit in its own vertex.
__global__ __launch_bounds__(256) void msbfs_init_kernel(int N, int s0, int ns, u64x4 *__restrict__ visited,
                                                         u64x4 *__restrict__ front)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        u64x4 m = {0ull, 0ull, 0ull, 0ull};
        const int64_t b = v - s0;
        if (b >= 0 && b < ns) m[b >> 6] = 1ull << (b & 63);
        visited[v] = m;
        front[v] = m;
    }
}

// One wave per vertex: next = (OR of the neighbours' frontier words) & ~visited.
// count += popcount(next) over the grid; done vertices (every source of the
// batch visited, `full`) only clear their next words.
__global__ __launch_bounds__(256) void msbfs_level_kernel(const int32_t *__restrict__ R,
                                                          const int32_t *__restrict__ C, int N,
                                                          const u64x4 *__restrict__ front,
                                                          u64x4 *__restrict__ next,
                                                          u64x4 *__restrict__ visited, u64x4 full,
                                                          unsigned long long *__restrict__ count)
{
    __shared__ unsigned long long wcnt[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t v = (int64_t)blockIdx.x * 4 + w;
    unsigned long long c = 0;
    if (v < N) {
        const u64x4 vis = visited[v];
        u64x4 acc = {0ull, 0ull, 0ull, 0ull};
        const bool done = ((vis & full)[0] == full[0]) && ((vis & full)[1] == full[1]) &&
                          ((vis & full)[2] == full[2]) && ((vis & full)[3] == full[3]);
        if (!done) {
            const int64_t e1 = R[v + 1];
            for (int64_t e = R[v] + lane; e < e1; e += 64) acc |= front[C[e]];
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] |= __shfl_xor(acc[q], d, 64);
            }
        }
        const u64x4 nx = acc & ~vis;
        if (lane == 0) {
            next[v] = nx;
            if (!done) visited[v] = vis | nx;
            c = (unsigned long long)(__popcll(nx[0]) + __popcll(nx[1]) + __popcll(nx[2]) + __popcll(nx[3]));
        }
    }
    if (lane == 0) wcnt[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        if (t) atomicAdd(count, t);
    }
}

hipError_t launch_msbfs_init(int N, int s0, int ns, void *visited, void *front, hipStream_t st)
{
    if (ns < 1 || ns > 256) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>(((int64_t)N + 255) / 256, 4096);
    hipLaunchKernelGGL(msbfs_init_kernel, dim3((unsigned)blocks), dim3(256), 0, st, N, s0, ns, (u64x4 *)visited,
                       (u64x4 *)front);
    return hipGetLastError();
}

hipError_t launch_msbfs_level(const int32_t *R, const int32_t *C, int N, const void *front, void *next,
                              void *visited, int ns, unsigned long long *count, hipStream_t st)
{
    if (ns < 1 || ns > 256) return hipErrorInvalidValue;
    u64x4 full = {0ull, 0ull, 0ull, 0ull};
    for (int b = 0; b < ns; b++) full[b >> 6] |= 1ull << (b & 63);
    hipLaunchKernelGGL(msbfs_level_kernel, dim3((unsigned)(((int64_t)N + 3) / 4)), dim3(256), 0, st, R, C, N,
                       (const u64x4 *)front, (u64x4 *)next, (u64x4 *)visited, full, count);
    return hipGetLastError();
}

// ---------------------------------------------------------------- BC
// Per-source state of a batch, each array N (lvl: N + 2) entries per source.
struct BcState {
    int32_t *dist, *queue, *lvl;
    double *sigma, *delta;
};

constexpr int BC_T = 256;

// One workgroup per source s0 + blockIdx.x: forward BFS counting shortest
// paths, then the dependencies level by level from the deepest.  out[b * N + v]
// gets delta_s(v) (0 for v = s and for vertices s does not reach).
__global__ __launch_bounds__(BC_T) void bc_source_kernel(const int32_t *__restrict__ R,
                                                         const int32_t *__restrict__ C, int N, int s0, int ns,
                                                         BcState st, double *__restrict__ out)
{
    const int b = blockIdx.x;
    if (b >= ns) return;
    const int s = s0 + b;
    const int64_t off = (int64_t)b * N;
    int32_t *dist = st.dist + off;
    int32_t *queue = st.queue + off;
    int32_t *lvl = st.lvl + (int64_t)b * (N + 2);
    double *sigma = st.sigma + off;
    double *delta = st.delta + off;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = BC_T / 64;
    __shared__ int s_tail;

    for (int v = tid; v < N; v += BC_T) {
        dist[v] = -1;
        sigma[v] = 0.0;
        delta[v] = 0.0;
    }
    __syncthreads();
    if (tid == 0) {
        dist[s] = 0;
        sigma[s] = 1.0;
        queue[0] = s;
        lvl[0] = 0;
        lvl[1] = 1;
        s_tail = 1;
    }
    __syncthreads();
    // forward: level l holds queue[lvl[l], lvl[l + 1])
    int l = 0, head = 0, tail = 1;
    while (head < tail) {
        for (int k = head + wave; k < tail; k += NW) {
            const int u = queue[k];
            const double su = sigma[u];
            const int e1 = R[u + 1];
            for (int e = R[u] + lane; e < e1; e += 64) {
                const int w = C[e];
                int d = __hip_atomic_load(&dist[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (d == -1) {
                    const int old = atomicCAS(&dist[w], -1, l + 1);
                    if (old == -1) queue[atomicAdd(&s_tail, 1)] = w;
                    d = old == -1 ? l + 1 : old;
                }
                if (d == l + 1) atomicAdd(&sigma[w], su);
            }
        }
        __syncthreads();
        head = tail;
        tail = s_tail;
        l++;
        if (tid == 0) lvl[l + 1] = tail;
        __syncthreads();
    }
    // backward: levels l - 1 .. 0 (the deepest level has no successors)
    for (int q = l - 1; q >= 0; q--) {
        const int a = lvl[q], z = lvl[q + 1];
        for (int k = a + wave; k < z; k += NW) {
            const int v = queue[k];
            const double sv = sigma[v];
            double part = 0.0;
            const int e1 = R[v + 1];
            for (int e = R[v] + lane; e < e1; e += 64) {
                const int w = C[e];
                if (dist[w] == q + 1) part += sv / sigma[w] * (1.0 + delta[w]);
            }
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d, 64);
            if (lane == 0) delta[v] = part;
        }
        __syncthreads();
    }
    for (int v = tid; v < N; v += BC_T) out[off + v] = v == s ? 0.0 : delta[v];
}

// bc[v] += sum over the batch's sources b (in order) of out[b * N + v].
__global__ __launch_bounds__(256) void bc_sum_kernel(const double *__restrict__ out, int N, int ns,
                                                     double *__restrict__ bc)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    double a = bc[v];
    for (int b = 0; b < ns; b++) a += out[(int64_t)b * N + v];
    bc[v] = a;
}

hipError_t launch_bc_batch(const int32_t *R, const int32_t *C, int N, int s0, int ns, int32_t *dist,
                           int32_t *queue, int32_t *lvl, double *sigma, double *delta, double *out, double *bc,
                           hipStream_t st)
{
    if (ns < 1 || N < 1) return hipErrorInvalidValue;
    const BcState s{dist, queue, lvl, sigma, delta};
    hipLaunchKernelGGL(bc_source_kernel, dim3((unsigned)ns), dim3(BC_T), 0, st, R, C, N, s0, ns, s, out);
    hipLaunchKernelGGL(bc_sum_kernel, dim3((unsigned)(((int64_t)N + 255) / 256)), dim3(256), 0, st, out, N, ns, bc);
    return hipGetLastError();
}

}  // namespace bna
