// k_graph.hip -- kernel (d): degree and eigenvector-centrality power iteration
// over a device CSR.
//
// Degree follows W/Degree/Degree/Degree.cpp:83-87: a float running sum of V in
// CSR order, so the order (and every rounding) is the reference's.  One lane per
// row adds its row in order from an LDS window that the wave fills by direct
// global-to-LDS copies; binary graphs take the row length
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

// One lane per row: the row's values are added one after another in CSR order,
// the reference's float loop (W/Degree/Degree/Degree.cpp:83-87), so the sum is
// the same float bit for bit.  Each row's sum is one chain of dependent adds,
// so what bounds the kernel is how many chains run side by side: a wave owns
// DEG_RPW consecutive rows (one lane each), so that the grid has several waves
// per SIMD to interleave.  The rows' values are one contiguous span of V; the
// whole wave copies it into an LDS window of DEG_WIN floats with
// global_load_lds (every lane moves 16 bytes per instruction, no registers in
// between), then the row lanes add their parts of the window.  A row longer
// than the window carries its running sum into the next window.
typedef __attribute__((address_space(1))) void deg_gvoid;
typedef __attribute__((address_space(3))) void deg_lvoid;
typedef float deg_f4 __attribute__((ext_vector_type(4)));
__device__ inline void deg_add4(float &s, const deg_f4 p)
{
    s += p[0];
    s += p[1];
    s += p[2];
    s += p[3];
}

constexpr int DEG_WIN = 4096;  // floats per window (16 KB of LDS per wave)
constexpr int DEG_RPW = 8;     // rows per wave

__global__ __launch_bounds__(64) void degree_kernel(const int32_t *__restrict__ R, const float *__restrict__ V, int N,
                                                    int64_t nnz, float *__restrict__ deg)
{
    __shared__ deg_f4 win[DEG_WIN / 4];
    const int lane = threadIdx.x;
    const int r0 = blockIdx.x * DEG_RPW;
    const int r = r0 + lane;
    const int rl = min(r0 + DEG_RPW, N);
    const bool mine = lane < DEG_RPW && r < N;
    const int b = mine ? R[r] : 0, e = mine ? R[r + 1] : 0;
    const int s0 = R[r0] & ~3;  // span start, down to a 16-byte boundary (V is 16-byte aligned)
    const int s1 = R[rl];
    // Every 16-byte piece copied lies inside V: a piece past the span rereads
    // the span's last piece, and none starts after the last whole piece of V
    // (vlast).  When V ends in a partial piece inside the span, its floats are
    // written into the window one by one.  Floats past s1 are never added.
    const int64_t vlast = ((nnz >> 2) << 2) - 4;
    const int64_t klast = min((int64_t)((s1 - 1) & ~3), vlast);
    float sum = 0.0f;
    for (int64_t w0 = s0; w0 < s1; w0 += DEG_WIN) {
        __syncthreads();  // the previous window's reads are done
#pragma unroll
        for (int j = 0; j < DEG_WIN / 256; j++) {
            const int64_t k = min(w0 + 4 * (64 * j + lane), klast);
            __builtin_amdgcn_global_load_lds((deg_gvoid *)(V + k), (deg_lvoid *)(win + 64 * j), 16, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0): the copies have landed (gfx9 encoding)
        if (s1 > vlast + 4 && lane == 0)
            for (int64_t k = max(vlast + 4, w0); k < min((int64_t)s1, w0 + DEG_WIN); k++)
                ((float *)win)[k - w0] = V[k];
        __syncthreads();
        // The lane's part of the window, [klo, khi) in floats, read as whole
        // 16-byte pieces: the first and last piece masked (a float outside the
        // part becomes -0.0, which leaves every sum unchanged), the pieces in
        // between added as they are, four pieces read ahead of the additions.
        // The additions are exactly the row's, in order.
        const int klo = (int)(max((int64_t)b, w0) - w0);
        const int khi = (int)(min((int64_t)e, w0 + DEG_WIN) - w0);
        if (klo < khi) {
            const int q0 = klo >> 2, q1 = (khi - 1) >> 2;  // first and last piece
            {
                const deg_f4 h = win[q0];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int k = 4 * q0 + c;
                    sum += (k >= klo && k < khi) ? h[c] : -0.0f;
                }
            }
            if (q1 > q0) {
                int q = q0 + 1;  // middle pieces [q0 + 1, q1)
                if (q + 4 <= q1) {
                    deg_f4 a0 = win[q], a1 = win[q + 1], a2 = win[q + 2], a3 = win[q + 3];
                    for (q += 4; q + 4 <= q1; q += 4) {
                        const deg_f4 b0 = win[q], b1 = win[q + 1], b2 = win[q + 2], b3 = win[q + 3];
                        deg_add4(sum, a0);
                        deg_add4(sum, a1);
                        deg_add4(sum, a2);
                        deg_add4(sum, a3);
                        a0 = b0;
                        a1 = b1;
                        a2 = b2;
                        a3 = b3;
                    }
                    deg_add4(sum, a0);
                    deg_add4(sum, a1);
                    deg_add4(sum, a2);
                    deg_add4(sum, a3);
                }
                for (; q < q1; q++) deg_add4(sum, win[q]);
                const deg_f4 t = win[q1];
#pragma unroll
                for (int c = 0; c < 4; c++) sum += (4 * q1 + c < khi) ? t[c] : -0.0f;
            }
        }
    }
    if (mine) deg[r] = sum;
}

// V with fewer than 4 floats (no whole 16-byte piece): one thread per row.
__global__ __launch_bounds__(256) void degree_small_kernel(const int32_t *__restrict__ R, const float *__restrict__ V,
                                                           int N, float *__restrict__ deg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float s = 0.0f;
    for (int k = R[i]; k < R[i + 1]; k++) s += V[k];
    deg[i] = s;
}

// Binary graphs: the row length as float (B/Degree/Degree/Degree.cpp:84-91).
__global__ __launch_bounds__(256) void degree_len_kernel(const int32_t *__restrict__ R, int N, float *__restrict__ deg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) deg[i] = (float)(R[i + 1] - R[i]);
}

hipError_t launch_degree(const int32_t *R, const float *V, int N, int64_t nnz, float *deg, hipStream_t st)
{
    if (N < 1) return hipSuccess;
    if (V && nnz >= 4)
        hipLaunchKernelGGL(degree_kernel, dim3((unsigned)((N + DEG_RPW - 1) / DEG_RPW)), dim3(64), 0, st, R, V, N, nnz,
                           deg);
    else if (V)
        hipLaunchKernelGGL(degree_small_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, R, V, N, deg);
    else
        hipLaunchKernelGGL(degree_len_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, R, N, deg);
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
    uint32_t arrive;  // workgroups of the running segmented SpMV that have finished
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
            // unconditional loads (a slot past the row end rereads the row's
            // last entry) so that they are all in flight together
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int tc = min(t0 + 64 * u, re - 1);
                cc[u] = C[tc];
                vv[u] = V ? V[tc] : 1.0f;
            }
            double xx[4];
#pragma unroll
            for (int u = 0; u < 4; u++) xx[u] = x[cc[u]];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double pr = (double)vv[u] * (xx[u] * inv);
                s += (t0 + 64 * u < re) ? pr : 0.0;
            }
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

// End of one iteration, from the sums of (y-x)^2, (y+x)^2 and y^2 over all rows:
// the stopping rule of weighted-net.cu:190-210 and the next 1/n.
__device__ inline void ec_advance(EcState *st, double sdm, double sdp, double syy)
{
    const double n_next = sqrt(syy);
    st->lambda = st->n_cur;
    st->inv_last = st->inv_n;
    st->err1 = sqrt(sdm);
    st->err2 = sqrt(sdp);
    st->iter += 1;
    if (!(st->err1 > st->tol && st->err2 > st->tol && st->iter < st->max_iter)) st->done = 1;
    st->n_cur = n_next;
    st->inv_n = 1.0 / n_next;
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
        ec_advance(st, sh[0][0], sh[1][0], sh[2][0]);
    }
}

// ---- segment-blocked SpMV (graphs whose rows average 16 or more entries)
// The plain kernel above gathers x[c] from L2 for every entry: each 8-byte
// gather moves a whole cache line to the CU, so it runs at a fraction of HBM
// speed.  Here the columns are cut into S segments of SEG voxels whose slice
// of x fits in LDS, and the rows into G blocks of RB rows, one workgroup each.
// Once per call the CSR is copied into a blocked layout: for block b, segment
// s and row i of the block, the row's entries with a column in segment s, in
// CSR order, padded to whole quads of 4 entries (column offset SEG, which
// reads the always-zero LDS slot past the segment, and value 0).
// A quad holds four 16-bit column offsets within the segment (8 bytes) and
// four values (16 bytes), so an entry costs 6 bytes instead of 8, and the
// quads of (b, s) are contiguous, so every iteration streams them in order.
// Per iteration workgroup b stages x[segment] * inv in LDS for each segment
// in turn and its 16-lane groups walk the rows' quads, gathering x from LDS.
// The row sums collect in LDS in a fixed order; after the rows are written,
// the last workgroup to finish reduces the per-workgroup partials in a fixed
// order and advances the iteration state, so an iteration is one launch.
constexpr int BLK_THREADS = 1024;
constexpr int BLK_LANES = 16;  // lanes per row
constexpr int BLK_UNROLL = 4;  // quads per lane in flight
constexpr int BLK_LDS_BUDGET = 160000;
constexpr int BLK_XS_PER_THREAD = (BLK_LDS_BUDGET / 8 + BLK_THREADS - 1) / BLK_THREADS;  // SEG <= budget / 8
constexpr int BLK_MIN_N = 4096;
constexpr int BLK_MAX_S = 64;
static_assert(BLK_MAX_S <= 64, "segment cursors live one per lane");

__device__ inline int wave_min_i(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ inline int wave_max_i(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Segment of a column; a column outside [0, N) (a malformed device CSR) is
// clamped to a segment so that nothing is written out of bounds.
__device__ inline int blk_seg(int c, int S, int SEG) { return min(max(c, 0) / SEG, S - 1); }

// qc[r S + s] = quads of row r in segment s.  One wave per row.
__global__ __launch_bounds__(256) void ec_blk_count_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                           int N, int S, int SEG, int32_t *__restrict__ qc)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + w;
    if (r >= N) return;
    const int rb = R[r], re = R[r + 1];
    int cnt = 0;  // lane s counts segment s
    for (int k0 = rb; k0 < re; k0 += 64) {
        const int k = k0 + lane;
        const int sg = k < re ? blk_seg(C[k], S, SEG) : -1;
        const int lo = wave_min_i(k < re ? sg : S), hi = wave_max_i(sg);
        for (int t = lo; t <= hi; t++) {
            const int m = __popcll(__ballot(sg == t));
            if (lane == t) cnt += m;
        }
    }
    if (lane < S) qc[(size_t)r * S + lane] = (cnt + 3) >> 2;
}

// Exclusive scan of the quad counts in (block, segment, row) order into
// qoff[0 .. items]; rows past N count 0.  Three steps: each workgroup scans
// its block's S RB items (local offsets, block total into btot[b]); one
// workgroup scans the G block totals into btot (and writes qoff[items]); each
// workgroup adds its block's base.
__device__ inline int32_t blk_qc(const int32_t *__restrict__ qc, int N, int RB, int S, int b, int k)
{
    const int sg = k / RB, i = k - sg * RB;
    const int64_t r = (int64_t)b * RB + i;
    return r < N ? qc[r * S + sg] : 0;
}

// Block-wide exclusive scan of one value per thread (1024 threads); returns
// the thread's exclusive prefix and the total through *tot.
__device__ inline int32_t blk_excl_scan(int32_t v, int32_t *sh, int32_t *tot)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int32_t u = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += u;
        __syncthreads();
    }
    const int32_t inc = sh[t];
    *tot = sh[1023];
    __syncthreads();
    return inc - v;
}

__global__ __launch_bounds__(1024) void ec_blk_scan_local_kernel(const int32_t *__restrict__ qc, int N, int RB,
                                                                 int S, int32_t *__restrict__ qoff,
                                                                 int32_t *__restrict__ btot)
{
    __shared__ int32_t sh[1024];
    const int b = blockIdx.x, t = threadIdx.x;
    const int nq = S * RB;
    const int per = (nq + 1023) / 1024;
    const int k0 = min(nq, t * per), k1 = min(nq, k0 + per);
    int32_t local = 0;
    for (int k = k0; k < k1; k++) local += blk_qc(qc, N, RB, S, b, k);
    int32_t tot;
    int32_t run = blk_excl_scan(local, sh, &tot);
    int32_t *q = qoff + (int64_t)b * nq;
    for (int k = k0; k < k1; k++) {
        q[k] = run;
        run += blk_qc(qc, N, RB, S, b, k);
    }
    if (t == 0) btot[b] = tot;
}

// G <= 1024 block totals -> exclusive bases in place; qoff[items] = total.
__global__ __launch_bounds__(1024) void ec_blk_scan_base_kernel(int G, int64_t items, int32_t *__restrict__ qoff,
                                                                int32_t *__restrict__ btot)
{
    __shared__ int32_t sh[1024];
    const int t = threadIdx.x;
    const int32_t v = t < G ? btot[t] : 0;
    int32_t tot;
    const int32_t base = blk_excl_scan(v, sh, &tot);
    if (t < G) btot[t] = base;
    if (t == 0) qoff[items] = tot;
}

__global__ __launch_bounds__(1024) void ec_blk_scan_add_kernel(int nq, int32_t *__restrict__ qoff,
                                                               const int32_t *__restrict__ btot)
{
    const int32_t base = btot[blockIdx.x];
    int32_t *q = qoff + (int64_t)blockIdx.x * nq;
    for (int k = threadIdx.x; k < nq; k += 1024) q[k] += base;
}

// Copies each row's entries to their quads, in CSR order within a segment,
// and zero-fills each slice up to its last quad.  One wave per row.
__global__ __launch_bounds__(256) void ec_blk_scatter_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                             const float *__restrict__ V, int N, int RB, int S, int SEG,
                                                             const int32_t *__restrict__ qoff, uint16_t *__restrict__ cq,
                                                             float *__restrict__ vq)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + w;
    if (r >= N) return;
    const int b = r / RB, i = r - b * RB;
    // lane s keeps segment s's next entry and its slice end
    int cur = 0, end = 0;
    if (lane < S) {
        const int64_t k = ((int64_t)b * S + lane) * RB + i;
        cur = 4 * qoff[k];
        end = 4 * qoff[k + 1];
    }
    const int rb = R[r], re = R[r + 1];
    for (int k0 = rb; k0 < re; k0 += 64) {
        const int k = k0 + lane;
        const bool act = k < re;
        const int c = act ? C[k] : 0;
        const float v = (act && V) ? V[k] : 0.0f;
        const int sg = act ? blk_seg(c, S, SEG) : -1;
        const int lo = wave_min_i(act ? sg : S), hi = wave_max_i(sg);
        for (int t = lo; t <= hi; t++) {
            const uint64_t m = __ballot(sg == t);
            const int base = __shfl(cur, t, 64);
            if (sg == t) {
                const int d = base + __popcll(m & ((1ull << lane) - 1ull));
                cq[d] = (uint16_t)min(max(c - t * SEG, 0), SEG - 1);
                if (V) vq[d] = v;
            }
            if (lane == t) cur += __popcll(m);
        }
    }
    if (lane < S) {
        for (int d = cur; d < end; d++) {
            cq[d] = (uint16_t)SEG;  // xs[SEG] is 0
            if (V) vq[d] = 0.0f;
        }
    }
}

template <bool WEIGHTED, int LANES = BLK_LANES, int UNROLL = BLK_UNROLL>
__global__ __launch_bounds__(BLK_THREADS) void spmv_ec_blk_kernel(const uint16_t *__restrict__ cq,
                                                                  const float *__restrict__ vq,
                                                                  const int32_t *__restrict__ qoff, int N, int RB,
                                                                  int S, int SEG, const double *__restrict__ x,
                                                                  double *__restrict__ y, double *__restrict__ part,
                                                                  EcState *__restrict__ st)
{
    if (st->done) return;
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    extern __shared__ double blk_lds[];
    double *xs = blk_lds;                 // SEG + 1 doubles: x * inv of the current segment, then 0
    double *acc = xs + SEG + 1;           // RB doubles: row sums
    int32_t *ql = (int32_t *)(acc + RB);  // S RB + 1 quad offsets of this block
    __shared__ double red[3][BLK_THREADS / 64];
    __shared__ int is_last;
    const int tid = threadIdx.x;
    const double inv = st->inv_n;
    const int r0 = blockIdx.x * RB;
    const int nr = min(RB, N - r0);
    const int nq = S * RB;
    {
        const int32_t *qg = qoff + (int64_t)blockIdx.x * nq;
        for (int i = tid; i <= nq; i += BLK_THREADS) ql[i] = qg[i];
        for (int i = tid; i < nr; i += BLK_THREADS) acc[i] = 0.0;
        if (tid == 0) xs[SEG] = 0.0;  // the padding slot; staging writes below SEG only
    }
    const int grp = tid / LANES, gl = tid % LANES;
    constexpr int NGRP = BLK_THREADS / LANES;
    const u2 *cq2 = (const u2 *)cq;
    const f4 *vq4 = (const f4 *)vq;
    for (int s = 0; s < S; s++) {
        const int c0 = s * SEG, n = min(SEG, N - c0);
        {
            // all of this thread's loads of the segment first, then the stores
            double t[BLK_XS_PER_THREAD];
#pragma unroll
            for (int j = 0; j < BLK_XS_PER_THREAD; j++) {
                const int i = tid + j * BLK_THREADS;
                t[j] = i < n ? x[c0 + i] : 0.0;
            }
            __syncthreads();  // the previous segment's gathers are done
#pragma unroll
            for (int j = 0; j < BLK_XS_PER_THREAD; j++) {
                const int i = tid + j * BLK_THREADS;
                if (i < n) xs[i] = t[j] * inv;
            }
        }
        __syncthreads();
        for (int ri = grp; ri < nr; ri += NGRP) {
            const int qs = ql[s * RB + ri], qe = ql[s * RB + ri + 1];
            double sum = 0.0;
            for (int q0 = qs + gl; q0 < qe; q0 += LANES * UNROLL) {
                // every load is issued (a slot past the slice rereads its
                // last quad) so they all go out together
                u2 cw[UNROLL];
                f4 vw[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    const int q = min(q0 + LANES * u, qe - 1);
                    cw[u] = cq2[q];
                    if (WEIGHTED) vw[u] = vq4[q];
                }
#pragma unroll
                for (int u = 0; u < UNROLL; u++) {
                    const bool in = q0 + LANES * u < qe;
                    const int o0 = cw[u][0] & 0xffff, o1 = cw[u][0] >> 16, o2 = cw[u][1] & 0xffff, o3 = cw[u][1] >> 16;
                    const double p0 = WEIGHTED ? (double)vw[u][0] * xs[o0] : xs[o0];
                    const double p1 = WEIGHTED ? (double)vw[u][1] * xs[o1] : xs[o1];
                    const double p2 = WEIGHTED ? (double)vw[u][2] * xs[o2] : xs[o2];
                    const double p3 = WEIGHTED ? (double)vw[u][3] * xs[o3] : xs[o3];
                    sum += in ? p0 : 0.0;
                    sum += in ? p1 : 0.0;
                    sum += in ? p2 : 0.0;
                    sum += in ? p3 : 0.0;
                }
            }
#pragma unroll
            for (int d = LANES / 2; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, LANES);
            if (gl == 0) acc[ri] += sum;
        }
    }
    __syncthreads();
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = tid; i < nr; i += BLK_THREADS) {
        const double yv = acc[i];
        y[r0 + i] = yv;
        const double xo = x[r0 + i];
        const double dm = yv - xo, dp = yv + xo;
        a += dm * dm;
        b += dp * dp;
        c += yv * yv;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
    }
    __syncthreads();
    if (tid < 3) {
        double t = 0.0;
        for (int k = 0; k < BLK_THREADS / 64; k++) t += red[tid][k];
        part[3 * blockIdx.x + tid] = t;
        __threadfence();
    }
    __syncthreads();
    if (tid == 0) is_last = (atomicAdd(&st->arrive, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!is_last) return;
    // the last workgroup: every other workgroup's partials are written
    __threadfence();
    a = b = c = 0.0;
    for (int k = tid; k < (int)gridDim.x; k += BLK_THREADS) {
        a += part[3 * k + 0];
        b += part[3 * k + 1];
        c += part[3 * k + 2];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    __syncthreads();
    if ((tid & 63) == 0) {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
    }
    __syncthreads();
    if (tid == 0) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < BLK_THREADS / 64; k++) {
            t0 += red[0][k];
            t1 += red[1][k];
            t2 += red[2][k];
        }
        st->arrive = 0;
        ec_advance(st, t0, t1, t2);
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
EcPlan ec_plan(int N, int64_t nnz, int ncu)
{
    EcPlan p{};
    if (N < BLK_MIN_N || nnz < 16 * (int64_t)N || ncu < 1 || ncu > 1024) return p;  // G <= ncu totals in one scan
    const int RB = (N + ncu - 1) / ncu;
    for (int S = 1; S <= BLK_MAX_S; S++) {
        const int SEG = (((N + S - 1) / S) + 7) / 8 * 8;
        if (SEG > 65535) continue;  // column offsets, and the padding offset SEG, are 16-bit
        const size_t lds =
            sizeof(double) * ((size_t)SEG + 1 + RB) + sizeof(int32_t) * ((size_t)S * RB + 1);
        if (lds > (size_t)BLK_LDS_BUDGET) continue;
        p.blocked = true;
        p.RB = RB;
        p.G = (N + RB - 1) / RB;
        p.S = S;
        p.SEG = SEG;
        p.lds = lds;
        p.items = (int64_t)p.G * S * RB;
        p.quads = (nnz + 3 * p.items) / 4 + 1;  // each slice pads fewer than 4 entries
        if (4 * p.quads >= (int64_t)INT32_MAX) return EcPlan{};  // entry indices are 32-bit
        return p;
    }
    return p;
}

hipError_t run_eigcent(const int32_t *R, const int32_t *C, const float *V, int N, const EcPlan &plan,
                       const EcBlocked &bl, double tol, int max_iter, int output, double *xa, double *xb, double *part,
                       void *state_dev, void *state_host, double *out, double *lambda, int *iters, bool use_graph,
                       const hipEvent_t ev[2], hipStream_t st)
{
    EcState *sd = (EcState *)state_dev;
    EcState *hs[2] = {(EcState *)state_host, (EcState *)((unsigned char *)state_host + EC_STATE_MAX_BYTES)};
    EcState &h = *hs[0];
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
    if (plan.blocked) {
        const unsigned row_waves = (unsigned)((N + 3) / 4);
        hipLaunchKernelGGL(ec_blk_count_kernel, dim3(row_waves), dim3(256), 0, st, R, C, N, plan.S, plan.SEG, bl.qc);
        const int nq = plan.S * plan.RB;
        hipLaunchKernelGGL(ec_blk_scan_local_kernel, dim3(plan.G), dim3(1024), 0, st, bl.qc, N, plan.RB, plan.S,
                           bl.qoff, bl.btot);
        hipLaunchKernelGGL(ec_blk_scan_base_kernel, dim3(1), dim3(1024), 0, st, plan.G, plan.items, bl.qoff, bl.btot);
        hipLaunchKernelGGL(ec_blk_scan_add_kernel, dim3(plan.G), dim3(1024), 0, st, nq, bl.qoff, bl.btot);
        hipLaunchKernelGGL(ec_blk_scatter_kernel, dim3(row_waves), dim3(256), 0, st, R, C, V, N, plan.RB, plan.S,
                           plan.SEG, bl.qoff, bl.cq, V ? bl.vq : nullptr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        e = hipFuncSetAttribute((const void *)(V ? spmv_ec_blk_kernel<true> : spmv_ec_blk_kernel<false>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)plan.lds);
        if (e != hipSuccess) return e;
    }
    double *buf[2] = {xa, xb};
    const int batch = 16;  // even, so every batch starts from buf[0]
    auto issue = [&](int k0, int k1) {
        for (int k = k0; k < k1; k++) {
            if (plan.blocked) {
                hipLaunchKernelGGL(V ? spmv_ec_blk_kernel<true> : spmv_ec_blk_kernel<false>, dim3(plan.G),
                                   dim3(BLK_THREADS), plan.lds, st, bl.cq, bl.vq, bl.qoff, N, plan.RB, plan.S,
                                   plan.SEG, buf[k & 1], buf[(k + 1) & 1], part, sd);
            } else {
                hipLaunchKernelGGL(spmv_ec_kernel, dim3(nb_spmv), dim3(256), 0, st, R, C, V, N, buf[k & 1],
                                   buf[(k + 1) & 1], part, sd);
                hipLaunchKernelGGL(ec_ctrl_kernel, dim3(1), dim3(1024), 0, st, part, nb_spmv, sd, 0);
            }
        }
    };
    // Whole batches replay one captured graph; the last, partial batch (and
    // every batch when graphs are off) is launched directly.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph && !plan.blocked && max_iter >= batch) {
        if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return e;
        issue(0, batch);
        e = hipStreamEndCapture(st, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
    }
    // Two batches in flight: batch b + 1 is queued before the host waits for
    // batch b's copy of the state, so the device never idles on the host.
    // Once the stopping rule fires, the kernels of a batch already queued
    // return at once; the state copied after batch b is final if it says done.
    int launched = 0, queued = 0, oldest = 0;
    int slot_of[2] = {0, 1};
    auto enqueue = [&](int slot) -> hipError_t {
        const int upto = (launched + batch < max_iter) ? launched + batch : max_iter;
        hipError_t r;
        if (exec && upto - launched == batch)
            r = hipGraphLaunch(exec, st);
        else {
            issue(launched, upto);
            r = hipGetLastError();
        }
        if (r != hipSuccess) return r;
        launched = upto;
        if ((r = hipMemcpyAsync(hs[slot], sd, sizeof(EcState), hipMemcpyDeviceToHost, st)) != hipSuccess) return r;
        return hipEventRecord(ev[slot], st);
    };
    const EcState *fin = nullptr;
    if (max_iter > 0) {
        e = enqueue(0);
        queued = 1;
        oldest = 0;
    }
    while (e == hipSuccess && queued > 0) {
        const int s_old = slot_of[oldest & 1];
        if (queued == 1 && launched < max_iter) {
            e = enqueue(s_old ^ 1);
            if (e != hipSuccess) break;
            slot_of[(oldest + 1) & 1] = s_old ^ 1;
            queued = 2;
        }
        if ((e = hipEventSynchronize(ev[s_old])) != hipSuccess) break;
        queued--;
        oldest++;
        if (hs[s_old]->done || queued == 0) {
            fin = hs[s_old];
            break;
        }
    }
    if (e == hipSuccess && queued > 0) e = hipStreamSynchronize(st);  // the batch behind the final one
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
    if (fin) h = *fin;
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
