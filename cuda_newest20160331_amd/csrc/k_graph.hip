// k_graph.hip -- kernel (d): degree and eigenvector-centrality power iteration
// over a device CSR.
//
// Degree follows W/Degree/Degree/Degree.cpp:83-87: a float running sum of V in
// CSR order, so the order (and every rounding) is the reference's.  One thread
// per row walks the row in order, in 16-byte pieces; binary graphs take the row length
// (B/Degree/Degree/Degree.cpp:84-91).
//
// Eigenvector centrality follows Lead_Vector_GPU (W/CUEC/weighted-net.cu:55-221):
//   x0 <- x; n = ||x||; x <- x * (1/n); y = A x; x <- y;
//   err1 = ||x - x0||; err2 = ||x + x0||; loop while err1 > tol && err2 > tol && it < max.
// One iteration is two launches: spmv_ec_kernel computes y = A (x * inv_n) with
// one wave per row (coalesced C and V reads, FP64 accumulation) over a fixed
// grid of EC_SPMV_BLOCKS workgroups that stride over the rows, and writes one
// partial sum of (y-x)^2, (y+x)^2 and y^2 per workgroup; ec_ctrl_kernel reduces
// the partials in a fixed order, applies the stopping rule, and stores the next
// 1/n.  Once the rule fires, every later launch returns at once, so the host
// queues iterations in batches of 16 and looks at the flag once per batch.  A
// batch is captured once into a HIP graph and replayed (32 launches become one).
#include "common.hpp"

namespace bna {

// One thread per row: the row's values are added one after another in CSR
// order, the reference's float loop (W/Degree/Degree/Degree.cpp:83-87), so the
// sum is the same float bit for bit.  A wave's 64 rows run side by side; each
// lane walks its own row in 16-byte pieces (after a scalar head up to the
// first 16-byte boundary), two pieces ahead of its additions, so a load
// instruction brings four values per lane instead of one.
__global__ __launch_bounds__(256) void degree_kernel(const int32_t *__restrict__ R, const float *__restrict__ V, int N,
                                                     float *__restrict__ deg)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int b = R[i], e = R[i + 1];
    if (!V) {
        deg[i] = (float)(e - b);
        return;
    }
    float s = 0.0f;
    int k = b;
    const int head = min(e, (b + 3) & ~3);  // V is 16-byte aligned
    for (; k < head; k++) s += V[k];
    for (; k + 8 <= e; k += 8) {
        const f4 p = *(const f4 *)(V + k), q = *(const f4 *)(V + k + 4);
        s += p[0];
        s += p[1];
        s += p[2];
        s += p[3];
        s += q[0];
        s += q[1];
        s += q[2];
        s += q[3];
    }
    if (k + 4 <= e) {
        const f4 p = *(const f4 *)(V + k);
        s += p[0];
        s += p[1];
        s += p[2];
        s += p[3];
        k += 4;
    }
    for (; k < e; k++) s += V[k];
    deg[i] = s;
}

hipError_t launch_degree(const int32_t *R, const float *V, int N, float *deg, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    hipLaunchKernelGGL(degree_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, R, V, N, deg);
    return hipGetLastError();
}

struct EcState {
    double inv_n;    // 1 / ||x|| for the coming iteration
    double n_cur;    // ||x|| for the coming iteration
    double lambda;   // n of the last finished iteration (the reference's vNorm)
    double inv_last; // 1/n of the last finished iteration
    double err1, err2;
    double tol;
    int32_t iter;
    int32_t max_iter;
    int32_t done;
    int32_t pad;
};

constexpr int EC_ROWS_PER_BLOCK = 4;  // waves per workgroup, one row at a time each
constexpr int EC_SPMV_BLOCKS = 16384;  // fixed grid, so the partial sums keep one order
static_assert(sizeof(EcState) <= EC_STATE_MAX_BYTES, "EcState outgrew EC_STATE_MAX_BYTES");
static_assert(EC_SPMV_BLOCKS <= EC_NPARTS_MAX && 4194304 / 256 <= EC_NPARTS_MAX, "ec_nparts bound");

__global__ __launch_bounds__(256) void ec_init_kernel(double *x, int N)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) x[i] = (double)i;  // init_AD, weighted-net.cu:38-43
}

// partial sums of x^2 per block, for the first norm
__global__ __launch_bounds__(256) void ec_norm_partial_kernel(const double *__restrict__ x, int N,
                                                              double *__restrict__ part)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = (i < N) ? x[i] : 0.0;
    sh[threadIdx.x] = v * v;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = 0.0;
        part[3 * blockIdx.x + 1] = 0.0;
        part[3 * blockIdx.x + 2] = sh[0];
    }
}

__global__ __launch_bounds__(256) void spmv_ec_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                      const float *__restrict__ V, int N,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      double *__restrict__ part, const EcState *__restrict__ st)
{
    if (st->done) return;
    __shared__ double sh[3][EC_ROWS_PER_BLOCK];
    const double inv = st->inv_n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int row = blockIdx.x * EC_ROWS_PER_BLOCK + w; row < N; row += gridDim.x * EC_ROWS_PER_BLOCK) {
        const int rb = R[row], re = R[row + 1];
        // four entries per lane in flight: their column and value loads are
        // issued together, then the four gathers of x, then the products
        double s = 0.0;
        for (int t0 = rb + lane; t0 < re; t0 += 4 * 64) {
            int cc[4];
            float vv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int t = t0 + 64 * u;
                cc[u] = t < re ? C[t] : -1;
                vv[u] = (t < re && V) ? V[t] : 1.0f;
            }
            double xx[4];
#pragma unroll
            for (int u = 0; u < 4; u++) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; u++) s += (double)vv[u] * (xx[u] * inv);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
        if (lane == 0) {
            y[row] = s;
            const double xo = x[row];
            const double dm = s - xo, dp = s + xo;
            a += dm * dm;
            b += dp * dp;
            c += s * s;
        }
    }
    if (lane == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
        sh[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int k = 0; k < EC_ROWS_PER_BLOCK; k++) t += sh[threadIdx.x][k];
        part[3 * blockIdx.x + threadIdx.x] = t;
    }
}

// Fixed-order reduction of nparts partial triples.  first = 1: only the norm of
// the start vector (no iteration is counted).
__global__ __launch_bounds__(1024) void ec_ctrl_kernel(const double *__restrict__ part, int nparts,
                                                       EcState *__restrict__ st, int first)
{
    if (!first && st->done) return;
    __shared__ double sh[3][1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int k = t; k < nparts; k += 1024) {
        a += part[3 * k + 0];
        b += part[3 * k + 1];
        c += part[3 * k + 2];
    }
    sh[0][t] = a;
    sh[1][t] = b;
    sh[2][t] = c;
    __syncthreads();
    for (int s = 512; s >= 1; s >>= 1) {
        if (t < s) {
            sh[0][t] += sh[0][t + s];
            sh[1][t] += sh[1][t + s];
            sh[2][t] += sh[2][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double n_next = sqrt(sh[2][0]);
        if (first) {
            st->n_cur = n_next;
            st->inv_n = 1.0 / n_next;
            st->iter = 0;
            st->done = (st->max_iter <= 0) ? 1 : 0;
            st->err1 = 1.0;
            st->err2 = 1.0;
            return;
        }
        st->lambda = st->n_cur;
        st->inv_last = st->inv_n;
        st->err1 = sqrt(sh[0][0]);
        st->err2 = sqrt(sh[1][0]);
        st->iter += 1;
        if (!(st->err1 > st->tol && st->err2 > st->tol && st->iter < st->max_iter)) st->done = 1;
        st->n_cur = n_next;
        st->inv_n = 1.0 / n_next;
    }
}

__global__ __launch_bounds__(256) void ec_scale_kernel(const double *__restrict__ x, int N,
                                                       const EcState *__restrict__ st, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] = x[i] * st->inv_last;
}

}  // namespace bna

// ---------------------------------------------------------------- host side
#include <vector>

namespace bna {

// Runs the whole power iteration on stream st.  Buffers: x0/x1 (N doubles),
// part (3 * nparts doubles), state (one EcState), out (N doubles, device);
// state_host: pinned host memory for one EcState (the state goes in and the
// stopping flag comes back through it, not through pageable staging).
// Returns the iteration count through *iters and the eigenvalue through *lambda.
hipError_t run_eigcent(const int32_t *R, const int32_t *C, const float *V, int N, double tol, int max_iter,
                       int output, double *xa, double *xb, double *part, void *state_dev, void *state_host,
                       double *out, double *lambda, int *iters, bool use_graph, hipStream_t st)
{
    EcState *sd = (EcState *)state_dev;
    EcState &h = *(EcState *)state_host;
    h = EcState{};
    h.tol = tol;
    h.max_iter = max_iter;
    hipError_t e = hipMemcpyAsync(sd, &h, sizeof(EcState), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    const int nb_norm = (N + 255) / 256;
    hipLaunchKernelGGL(ec_init_kernel, dim3(nb_norm), dim3(256), 0, st, xa, N);
    hipLaunchKernelGGL(ec_norm_partial_kernel, dim3(nb_norm), dim3(256), 0, st, xa, N, part);
    hipLaunchKernelGGL(ec_ctrl_kernel, dim3(1), dim3(1024), 0, st, part, nb_norm, sd, 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;

    const int rows_blocks = (N + EC_ROWS_PER_BLOCK - 1) / EC_ROWS_PER_BLOCK;
    const int nb_spmv = rows_blocks < EC_SPMV_BLOCKS ? rows_blocks : EC_SPMV_BLOCKS;
    double *buf[2] = {xa, xb};
    const int batch = 16;  // even, so every batch starts from buf[0]
    auto issue = [&](int k0, int k1) {
        for (int k = k0; k < k1; k++) {
            hipLaunchKernelGGL(spmv_ec_kernel, dim3(nb_spmv), dim3(256), 0, st, R, C, V, N, buf[k & 1],
                               buf[(k + 1) & 1], part, sd);
            hipLaunchKernelGGL(ec_ctrl_kernel, dim3(1), dim3(1024), 0, st, part, nb_spmv, sd, 0);
        }
    };
    // Whole batches replay one captured graph; the last, partial batch (and
    // every batch when graphs are off) is launched directly.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph && max_iter >= batch) {
        if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
        issue(0, batch);
        e = hipStreamEndCapture(st, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
    }
    int launched = 0;
    while (launched < max_iter) {
        const int upto = (launched + batch < max_iter) ? launched + batch : max_iter;
        if (exec && upto - launched == batch)
            e = hipGraphLaunch(exec, st);
        else {
            issue(launched, upto);
            e = hipGetLastError();
        }
        if (e != hipSuccess) break;
        launched = upto;
        if ((e = hipMemcpyAsync(&h, sd, sizeof(EcState), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) break;
        if (h.done) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
    const int K = h.iter;  // iterations actually executed
    if (K == 0) {
        // no iteration: the reference returns x as initialised and n = 0
        e = hipMemcpyAsync(out, xa, sizeof(double) * N, hipMemcpyDeviceToDevice, st);
        *lambda = 0.0;
    } else if (output == 1) {
        hipLaunchKernelGGL(ec_scale_kernel, dim3(nb_norm), dim3(256), 0, st, buf[(K - 1) & 1], N, sd, out);
        e = hipGetLastError();
        *lambda = h.lambda;
    } else {
        e = hipMemcpyAsync(out, buf[K & 1], sizeof(double) * N, hipMemcpyDeviceToDevice, st);
        *lambda = h.lambda;
    }
    if (iters) *iters = K;
    return e;
}

size_t ec_state_bytes() { return sizeof(EcState); }
int ec_nparts(int N) {
    int a = (N + EC_ROWS_PER_BLOCK - 1) / EC_ROWS_PER_BLOCK;
    a = a < EC_SPMV_BLOCKS ? a : EC_SPMV_BLOCKS;
    const int b = (N + 255) / 256;
    return a > b ? a : b;
}

}  // namespace bna
