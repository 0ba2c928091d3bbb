// k_csr.hip -- FP64 recheck of band pairs and the second half of kernel (c):
// row lengths, the exclusive scan into R, and the placement of the per-tile
// lists into CSR rows.
//
// CSR layout follows FormCSRGraph (W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:889-918):
// row i lists every neighbour j != i in ascending order with V = r(i, j).
// Row i's neighbours come in nb + 1 fragments: fragment k <= I(i) is the
// column-view list of tile (k, I(i)) for column i (neighbours in tile-row k,
// ascending), fragment k > I(i) is the row-view list of tile (I(i), k - 1) for
// row i (neighbours in tile-column k - 1, ascending).  Concatenating the
// fragments in k order is therefore already column-sorted, and the placement
// is a copy with the recheck's rejected entries squeezed out.
#include "common.hpp"

#include <algorithm>

namespace bna {

// One wave per candidate.  z[l] = (x[l] - mean) / norm is recomputed in double,
// the expression the normalisation evaluates, so the values equal it bit for
// bit.  r64 is evaluated in the FP64 oracle's order: lane q
// accumulates l = q, q + 64, ... with fused multiply-add, then a halving tree
// (lane q adds lane q + s for s = 32, 16, ..., 1).  Edge iff r64 > thr.
// stats[0]: max |r_mfma - r64| as double bits, stats[1]: decisions that differ
// from the MFMA value's own (r > thr) decision.
__global__ __launch_bounds__(256) void recheck_kernel(const float *__restrict__ xs, int64_t ld,
                                                      const double *__restrict__ mu,
                                                      const double *__restrict__ sd, int N, int L, int nb,
                                                      const Candidate *__restrict__ cands,
                                                      const unsigned long long *__restrict__ cand_count,
                                                      int64_t cand_cap, double thr,
                                                      int32_t *__restrict__ frag_live, Entry *__restrict__ entries,
                                                      int64_t entry_cap, unsigned long long *__restrict__ stats,
                                                      int mode, double *__restrict__ r64_out)
{
    const int64_t n = (int64_t)min((unsigned long long)cand_cap, *cand_count);
    const int lane = threadIdx.x & 63;
    double wmax = 0.0;       // this wave's max |r_mfma - r64| (lane 0)
    unsigned long long wflips = 0;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t c = wave; c < n; c += nwaves) {
        const Candidate cd = cands[c];
        if (cd.slot_row >= entry_cap || cd.slot_col >= entry_cap) continue;  // lists overflowed: rerun follows
        if (mode == 2) {
            if (lane == 0) r64_out[c] = (double)cd.r;
            continue;
        }
        const float *xi = xs + (int64_t)cd.i * ld;
        const float *xj = xs + (int64_t)cd.j * ld;
        const double mi = mu[cd.i], si = sd[cd.i], mj = mu[cd.j], sj = sd[cd.j];
        double p = 0.0;
#pragma unroll 4
        for (int l = lane; l < L; l += 64) {
            const double zi = ((double)xi[l] - mi) / si;
            const double zj = ((double)xj[l] - mj) / sj;
            p = __builtin_fma(zi, zj, p);
        }
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
            double o = __shfl_down(p, s, 64);
            if (lane < s) p = p + o;
        }
        if (lane == 0 && mode == 1) {
            r64_out[c] = p;
            entries[cd.slot_row].r = (float)p;
            entries[cd.slot_col].r = (float)p;
            wmax = fmax(wmax, fabs((double)cd.r - p));
        } else if (lane == 0) {
            const bool edge = p > thr;
            const bool raw = (double)cd.r > thr;
            if (edge) {
                entries[cd.slot_row].r = (float)p;
                entries[cd.slot_col].r = (float)p;
            } else {
                entries[cd.slot_row].col = -1;
                entries[cd.slot_col].col = -1;
                atomicSub(&frag_live[(int64_t)((cd.j / TILE) + 1) * N + cd.i], 1);
                atomicSub(&frag_live[(int64_t)(cd.i / TILE) * N + cd.j], 1);
            }
            wmax = fmax(wmax, fabs((double)cd.r - p));
            wflips += (edge != raw);
        }
    }
    if (lane == 0 && (wmax > 0.0 || wflips)) {
        atomicMax(&stats[0], (unsigned long long)__builtin_bit_cast(unsigned long long, wmax));
        if (wflips) atomicAdd(&stats[1], wflips);
    }
}

// Fragment tables are written fragment-major ([k][row]) by the tile kernel,
// where each tile's writes are contiguous.  This pass turns them row-major
// ([row][k]) through LDS, 64 x 64 at a time, so the placement reads a row's
// fragments contiguously, and adds each row's live counts into rowlen (which
// must start at zero; integer atomics, so the sum is exact).  A thread issues
// its loads of all three tables before the one barrier, so they are in flight
// together.
constexpr int TR = 64;
__global__ __launch_bounds__(256) void frag_rows_kernel(const int32_t *__restrict__ orig_km,
                                                        const int32_t *__restrict__ live_km,
                                                        const int64_t *__restrict__ off_km, int N, int nb1,
                                                        int32_t *__restrict__ orig_rm, int32_t *__restrict__ live_rm,
                                                        int64_t *__restrict__ off_rm,
                                                        unsigned long long *__restrict__ rowlen)
{
    __shared__ int64_t t64[TR][TR + 1];
    __shared__ int32_t to[TR][TR + 1], tl[TR][TR + 1];
    const int k0 = blockIdx.x * TR, r0 = blockIdx.y * TR;
    const int tx = threadIdx.x & (TR - 1), ty = threadIdx.x / TR;  // 64 x 4
    int64_t a[TR / 4];
    int32_t b[TR / 4], c[TR / 4];
#pragma unroll
    for (int u = 0; u < TR / 4; u++) {
        const int gk = k0 + ty + 4 * u, gr = r0 + tx;
        const bool in = gk < nb1 && gr < N;
        const int64_t e = (int64_t)gk * N + gr;
        a[u] = in ? off_km[e] : 0;
        b[u] = in ? orig_km[e] : 0;
        c[u] = in ? live_km[e] : 0;
    }
#pragma unroll
    for (int u = 0; u < TR / 4; u++) {
        const int k = ty + 4 * u;
        t64[k][tx] = a[u];
        to[k][tx] = b[u];
        tl[k][tx] = c[u];
    }
    __syncthreads();
    for (int r = ty; r < TR; r += 4) {
        const int gr = r0 + r, gk = k0 + tx;
        const int v = tl[tx][r];
        if (gr < N && gk < nb1) {
            const int64_t o = (int64_t)gr * nb1 + gk;
            off_rm[o] = t64[tx][r];
            orig_rm[o] = to[tx][r];
            live_rm[o] = v;
        }
        int sum = v;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
        if (tx == 0 && gr < N && sum) atomicAdd(&rowlen[gr], (unsigned long long)sum);
    }
}

// Exclusive scan of the N row lengths into R, in three passes: per-block sums
// (SCAN_E elements per block), a single-block scan of those sums, then each
// block scans its own elements from its offset.  R[N] is the total.
constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 8;
constexpr int SCAN_E = SCAN_T * SCAN_PER;

__device__ __forceinline__ int64_t block_scan_excl64(int64_t v, int64_t *sh, int64_t &total)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int64_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    int64_t pre = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < SCAN_T / 64; k++) {
        const int64_t x = sh[k];
        if (k < w) pre += x;
        total += x;
    }
    __syncthreads();
    return pre + inc - v;
}

__global__ __launch_bounds__(SCAN_T) void scan_sums_kernel(const int64_t *__restrict__ rowlen, int N,
                                                           int64_t *__restrict__ bsum)
{
    __shared__ int64_t sh[SCAN_T / 64];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_E;
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int64_t k = b0 + q * SCAN_T + threadIdx.x;
        if (k < N) s += rowlen[k];
    }
    int64_t total;
    block_scan_excl64(s, sh, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_T) void scan_top_kernel(int64_t *__restrict__ bsum, int nblocks)
{
    __shared__ int64_t sh[SCAN_T / 64];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += SCAN_T) {
        const int k = b0 + threadIdx.x;
        const int64_t v = k < nblocks ? bsum[k] : 0;
        int64_t total;
        const int64_t ex = block_scan_excl64(v, sh, total);
        if (k < nblocks) bsum[k] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nblocks] = carry;
}

__global__ __launch_bounds__(SCAN_T) void scan_down_kernel(const int64_t *__restrict__ rowlen, int N,
                                                           const int64_t *__restrict__ bsum, int nblocks,
                                                           int64_t *__restrict__ R64, int32_t *__restrict__ R32)
{
    __shared__ int64_t sh[SCAN_T / 64];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_E;
    // thread t owns elements b0 + t * SCAN_PER .. + SCAN_PER - 1 (contiguous run)
    int64_t v[SCAN_PER];
    int64_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int64_t k = b0 + (int64_t)threadIdx.x * SCAN_PER + q;
        v[q] = k < N ? rowlen[k] : 0;
        s += v[q];
    }
    int64_t total;
    int64_t run = bsum[blockIdx.x] + block_scan_excl64(s, sh, total);
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const int64_t k = b0 + (int64_t)threadIdx.x * SCAN_PER + q;
        if (k < N) {
            R64[k] = run;
            if (R32) R32[k] = (int32_t)run;
        }
        run += v[q];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        R64[N] = bsum[nblocks];
        if (R32) R32[N] = (int32_t)bsum[nblocks];
    }
}

// One wave per row: lane k takes fragment k (in chunks of 64) of the row-major
// tables, a wave scan of the live counts gives each fragment its destination,
// and the lane copies the fragment's entries, skipping rejected ones.
// Fragment k of row i lists neighbours in tile k (k <= I(i)) or tile k - 1
// (k > I(i)), each ascending, so the copies come out column-sorted.
__global__ __launch_bounds__(256) void place_kernel(const int32_t *__restrict__ frag_orig,
                                                    const int32_t *__restrict__ frag_live,
                                                    const int64_t *__restrict__ frag_off,
                                                    const Entry *__restrict__ entries,
                                                    const int64_t *__restrict__ R64, int N, int nb,
                                                    int32_t *__restrict__ Cout, float *__restrict__ Vout)
{
    const int lane = threadIdx.x & 63;
    const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (row >= N) return;
    const int nb1 = nb + 1;
    int64_t run = R64[row];
    for (int k0 = 0; k0 < nb1; k0 += 64) {
        const int k = k0 + lane;
        int live = 0, orig = 0;
        int64_t off = 0;
        if (k < nb1) {
            const int64_t f = row * nb1 + k;
            orig = frag_orig[f];
            if (orig) {
                live = frag_live[f];
                off = frag_off[f];
            }
        }
        int inc = live;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int t = __shfl_up(inc, d, 64);
            if (lane >= d) inc += t;
        }
        const int tot = __shfl(inc, 63, 64);
        int64_t dst = run + inc - live;
        if (live == orig) {
            for (int e = 0; e < orig; e++) {
                const Entry en = entries[off + e];
                Cout[dst + e] = en.col;
                if (Vout) Vout[dst + e] = en.r;
            }
        } else {
            for (int e = 0; e < orig; e++) {
                const Entry en = entries[off + e];
                if (en.col >= 0) {
                    Cout[dst] = en.col;
                    if (Vout) Vout[dst] = en.r;
                    dst++;
                }
            }
        }
        run += tot;
    }
}

// Value refinement of a weighted CSR: every kept pair's value is recomputed as
// V = (float) sum_l z_i[l] z_j[l] with z = (x - mean) * (1 / norm) in double,
// so |V - r64| stays within half a float ulp of r plus ~1e-15, whatever the MFMA
// accumulation did.  In the general form below one wave per row i takes the
// entries j > i, four at a time: each 16-lane group owns one entry, reads row j
// in 16-byte pieces (row i's pieces come from cache), sums in double and writes
// the value to entry (i, j) and to its mirror (j, i), which the group finds in
// row j's sorted columns.
constexpr int RG = 16;  // lanes per entry

__device__ __forceinline__ int64_t group_find_col(const int32_t *__restrict__ C, int64_t lo, int64_t hi, int key,
                                                  int gl, uint64_t gmask, int gbase, int N = 0)
{
    // first probe (N > 0): the RG entries around where key would sit if the
    // row's columns were spread evenly over [0, N); a row's columns ascend, so
    // the probe either holds key or tells on which side it lies
    if (N > 0 && hi - lo > RG) {
        int64_t p0 = lo + (int64_t)((double)(hi - lo) * (double)key / (double)N) - RG / 2;
        p0 = p0 < lo ? lo : (p0 > hi - RG ? hi - RG : p0);
        const int c = C[p0 + gl];
        const uint64_t eq = (__ballot(c == key) & gmask) >> gbase;
        if (eq) return p0 + __builtin_ctzll(eq);
        const uint64_t le = (__ballot(c < key) & gmask) >> gbase;
        if (le == 0) hi = p0;                      // key is below the probe
        else if (le == 0xffffull) lo = p0 + RG;    // key is above it
        else return -1;                            // between two neighbours: absent
    }
    // narrow [lo, hi) by sampling RG evenly spaced entries until one load covers it
    while (hi - lo > RG) {
        const int64_t step = (hi - lo + RG - 1) / RG;
        const int64_t p = lo + (int64_t)gl * step;
        const bool le = p < hi && C[p] <= key;
        const uint64_t m = (__ballot(le) & gmask) >> gbase;
        const int t = m ? 63 - __builtin_clzll(m) : 0;
        lo = lo + (int64_t)t * step;
        hi = min(hi, lo + step);
    }
    const int64_t p = lo + gl;
    const uint64_t m = (__ballot(p < hi && C[p] == key) & gmask) >> gbase;
    return m ? lo + __builtin_ctzll(m) : -1;
}

template <bool VEC>
__global__ __launch_bounds__(256) void refine_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                     float *__restrict__ V, const float *__restrict__ xs,
                                                     int64_t ld, const double *__restrict__ mu,
                                                     const double *__restrict__ sd, int N, int L, int row0, int row1)
{
    const int lane = threadIdx.x & 63;
    const int g = lane / RG, gl = lane % RG, gbase = g * RG;
    const uint64_t gmask = 0xffffull << gbase;
    const int64_t i = row0 + (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (i >= row1) return;
    const int64_t r0 = R[i], r1 = R[i + 1];
    if (r0 == r1) return;
    // first entry with column > i
    int64_t e0 = r1;
    for (int64_t b = r0; b < r1; b += 64) {
        const uint64_t m = __ballot(b + lane < r1 && C[b + lane] > (int)i);
        if (m) {
            e0 = b + __builtin_ctzll(m);
            break;
        }
    }
    const float *xi = xs + i * ld;
    const double mi = mu[i], ii = 1.0 / sd[i];
    // every group runs the same number of trips so the ballots stay wave-wide
    const int64_t trips = (r1 - e0 + 3) / 4;
    for (int64_t t = 0; t < trips; t++) {
        const int64_t e = e0 + 4 * t + g;
        const bool act = e < r1;
        const int j = act ? C[e] : (int)i;
        const float *xj = xs + (int64_t)j * ld;
        const double mj = mu[j], ij = 1.0 / sd[j];
        double p = 0.0;
        if (act) {
            if constexpr (VEC) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                const int L4 = L >> 2;
#pragma unroll 4
                for (int q = gl; q < L4; q += RG) {
                    const f4 a = *(const f4 *)(xi + 4 * q);
                    const f4 b = *(const f4 *)(xj + 4 * q);
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        p = __builtin_fma(((double)a[c] - mi) * ii, ((double)b[c] - mj) * ij, p);
                }
                for (int l = 4 * L4 + gl; l < L; l += RG)
                    p = __builtin_fma(((double)xi[l] - mi) * ii, ((double)xj[l] - mj) * ij, p);
            } else {
#pragma unroll 4
                for (int l = gl; l < L; l += RG)
                    p = __builtin_fma(((double)xi[l] - mi) * ii, ((double)xj[l] - mj) * ij, p);
            }
        }
#pragma unroll
        for (int s = RG / 2; s >= 1; s >>= 1) p += __shfl_xor(p, s, 64);
        const int64_t q = act ? group_find_col(C, R[j], R[j + 1], (int)i, gl, gmask, gbase, N)
                              : group_find_col(C, 0, 0, 0, gl, gmask, gbase);
        if (act && gl == 0) {
            const float v = (float)p;
            V[e] = v;
            if (q >= 0) V[q] = v;
        }
    }
}

// Chunked, row-resident form of the refinement for 16-byte aligned rows of at
// most 256*T floats.  The CSR entries of rows [row0, row1) are cut into chunks of
// REFINE_CHUNK entries and each wave takes chunks in turn, so a row with
// thousands of neighbours is spread over many waves instead of keeping one busy
// after the others are done.  For each row segment of its chunk the wave keeps
// row i as double z_i = (x_i - mean_i) / norm_i in registers (T pieces of four
// per lane), so only the rows j are read: four entries at a time, each as
// 64-lane 16-byte pieces, all T*4 loads issued before any is used.  The sum is
// taken over the raw x_j,
//   r = (sum_l z_i[l] x_j[l] - mean_j sum_l z_i[l]) / norm_j,
// which equals sum_l z_i[l] z_j[l] and needs one conversion and one FMA per
// element.  The four sums are folded across the wave with a butterfly that
// leaves entry b's total in lanes 16b..16b+15, the group that then finds and
// writes its mirror.  Every double sum has relative error below 1e-13, so the
// stored float differs from r64 by its own rounding only.  Only entries with
// j > i are computed; their mirrors (j, i) are written with them.
constexpr int REFINE_CHUNK = 128;

template <int T>
__global__ __launch_bounds__(256) void refine_rows_kernel(const int32_t *__restrict__ R, const int32_t *__restrict__ C,
                                                          float *__restrict__ V, const float *__restrict__ xs,
                                                          int64_t ld, const double *__restrict__ mu,
                                                          const double *__restrict__ sd, int N, int L, int row0,
                                                          int row1)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int EB = 4;  // entries per trip, one per 16-lane group at the end
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, gl = lane & 15, gbase = g * 16;
    const uint64_t gmask = 0xffffull << gbase;
    const bool h5 = lane & 32, h4 = lane & 16;
    const int L4 = L >> 2, Lt = L & 3;
    const int64_t base = R[row0], end = R[row1];
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t w0 = __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    for (int64_t w = w0;; w += nwaves) {
        const int64_t a = base + w * REFINE_CHUNK;
        if (a >= end) break;
        const int64_t b = min(end, a + REFINE_CHUNK);
        // the row of entry a: R[lo] <= a < R[hi], hi = lo + 1
        int lo = row0, hi = row1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if ((int64_t)R[mid] <= a) lo = mid;
            else hi = mid;
        }
        int64_t s0 = a;
        for (int64_t i = lo; s0 < b; i++) {
            const int64_t s1 = min(b, (int64_t)R[i + 1]);
            // first entry of the segment with column > i (a row's columns ascend)
            int64_t e0 = s1;
            for (int64_t bb = s0; bb < s1; bb += 64) {
                const uint64_t m = __ballot(bb + lane < s1 && C[bb + lane] > (int)i);
                if (m) {
                    e0 = bb + __builtin_ctzll(m);
                    break;
                }
            }
            s0 = s1;
            if (e0 == s1) continue;
            const float *xi = xs + i * ld;
            const double mi = mu[i], ii = 1.0 / sd[i];
            double zi[T][4], zt = 0.0, s = 0.0;
#pragma unroll
            for (int t = 0; t < T; t++) {
                const int q = lane + 64 * t;
                f4 x4 = {0.f, 0.f, 0.f, 0.f};
                if (q < L4) x4 = *(const f4 *)(xi + 4 * q);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    zi[t][c] = q < L4 ? ((double)x4[c] - mi) * ii : 0.0;
                    s += zi[t][c];
                }
            }
            if (lane < Lt) {
                zt = ((double)xi[4 * L4 + lane] - mi) * ii;
                s += zt;
            }
#pragma unroll
            for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k, 64);
            for (int64_t e = e0; e < s1; e += EB) {
                int j[EB];
                f4 xj[EB][T];
#pragma unroll
                for (int q4 = 0; q4 < EB; q4++) {
                    j[q4] = e + q4 < s1 ? C[e + q4] : -1;
                    const float *xr = xs + (int64_t)max(j[q4], 0) * ld;
#pragma unroll
                    for (int t = 0; t < T; t++) {
                        const int q = lane + 64 * t;
                        xj[q4][t] = (f4){0.f, 0.f, 0.f, 0.f};
                        if (j[q4] >= 0 && q < L4) xj[q4][t] = *(const f4 *)(xr + 4 * q);
                    }
                }
                double acc[EB];
#pragma unroll
                for (int q4 = 0; q4 < EB; q4++) {
                    double av = 0.0;
                    if (lane < Lt && j[q4] >= 0) av = zt * (double)xs[(int64_t)j[q4] * ld + 4 * L4 + lane];
#pragma unroll
                    for (int t = 0; t < T; t++)
#pragma unroll
                        for (int c = 0; c < 4; c++) av = __builtin_fma(zi[t][c], (double)xj[q4][t][c], av);
                    acc[q4] = av;
                }
                // fold: after xor 32 lane half h5 holds entries {2h5, 2h5+1}; after xor
                // 16 quarter (h5, h4) holds entry 2h5 + h4; xor 8..1 finish the sum
                double v0 = (h5 ? acc[2] : acc[0]) + __shfl_xor(h5 ? acc[0] : acc[2], 32, 64);
                double v1 = (h5 ? acc[3] : acc[1]) + __shfl_xor(h5 ? acc[1] : acc[3], 32, 64);
                double wv = (h4 ? v1 : v0) + __shfl_xor(h4 ? v0 : v1, 16, 64);
#pragma unroll
                for (int k = 8; k >= 1; k >>= 1) wv += __shfl_xor(wv, k, 64);
                const int jg = g == 0 ? j[0] : g == 1 ? j[1] : g == 2 ? j[2] : j[3];
                const int64_t q = jg >= 0 ? group_find_col(C, R[jg], R[jg + 1], (int)i, gl, gmask, gbase, N)
                                          : group_find_col(C, 0, 0, 0, gl, gmask, gbase);
                if (jg >= 0 && gl == 0) {
                    const float v = (float)((wv - mu[jg] * s) / sd[jg]);
                    V[e + g] = v;
                    if (q >= 0) V[q] = v;
                }
            }
        }
    }
}

// Value refinement on the tile lists, before placement.  One wave per row i
// keeps z_i in registers as double (T pieces of four per lane, as above) and
// takes the entries of its row-view fragments (neighbours j > i, tiles
// (I(i), J >= I(i))), four at a time, one per 16-lane group at the end: row j
// is read as 64-lane 16-byte pieces and
//   r = (sum_l z_i[l] x_j[l] - mean_j sum_l z_i[l]) / norm_j.
// The value goes to the entry and to the pair's column-view entry, whose slot
// is the column's fragment offset in tile (I(i), J(j)) plus the position the
// tile kernel recorded for the pair (mirror), so no search is needed.
// Rejected entries (column -1) are skipped.  A row's entry slots are queued in
// LDS first, so sparse fragments fill whole four-entry trips.  Workgroups are
// single waves: rows differ in length, and a wave that finishes frees its
// slot at once instead of waiting for the longest row of its workgroup.
constexpr int RQ = 512;  // queued slots per wave (a fragment holds at most TILE)
static_assert(RQ >= TILE, "a whole fragment fits the queue");

template <int T>
__global__ __launch_bounds__(64) void refine_lists_kernel(Entry *__restrict__ entries,
                                                           const uint8_t *__restrict__ mirror,
                                                           const int32_t *__restrict__ orig_rm,
                                                           const int64_t *__restrict__ off_rm,
                                                           const int64_t *__restrict__ off_km,
                                                           const float *__restrict__ xs, int64_t ld,
                                                           const double *__restrict__ mu,
                                                           const double *__restrict__ sd, int N, int L, int nb,
                                                           int row0, int row1)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    __shared__ int64_t q[RQ];
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, gl = lane & 15;
    const bool h5 = lane & 32, h4 = lane & 16;
    const int L4 = L >> 2, Lt = L & 3;
    const int64_t i = (int64_t)row0 + blockIdx.x;
    const int nb1 = nb + 1, I = (int)(i / TILE);
    float *V = (float *)entries;  // entry e's value is float 2e + 1

    const float *xi = xs + i * ld;
    const double mi = mu[i], ii = 1.0 / sd[i];
    double zi[T][4], zt = 0.0, s = 0.0;
#pragma unroll
    for (int t = 0; t < T; t++) {
        const int pq = lane + 64 * t;
        f4 x4 = {0.f, 0.f, 0.f, 0.f};
        if (pq < L4) x4 = *(const f4 *)(xi + 4 * pq);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            zi[t][c] = pq < L4 ? ((double)x4[c] - mi) * ii : 0.0;
            s += zi[t][c];
        }
    }
    if (lane < Lt) {
        zt = ((double)xi[4 * L4 + lane] - mi) * ii;
        s += zt;
    }
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k, 64);

    // refines the queued entries q[0, nq)
    auto flush = [&](int nq) {
        for (int e0 = 0; e0 < nq; e0 += 4) {
            int j[4];
            int64_t sl[4];
            f4 xj[4][T];
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) {
                sl[q4] = e0 + q4 < nq ? q[e0 + q4] : -1;
                j[q4] = sl[q4] >= 0 ? entries[sl[q4]].col : -1;
                const float *xr = xs + (int64_t)max(j[q4], 0) * ld;
#pragma unroll
                for (int t = 0; t < T; t++) {
                    const int pq = lane + 64 * t;
                    xj[q4][t] = (f4){0.f, 0.f, 0.f, 0.f};
                    if (j[q4] >= 0 && pq < L4) xj[q4][t] = *(const f4 *)(xr + 4 * pq);
                }
            }
            // this group's entry and its mirror slot, loaded beside the rows
            const int jg = g == 0 ? j[0] : g == 1 ? j[1] : g == 2 ? j[2] : j[3];
            const int64_t sg = g == 0 ? sl[0] : g == 1 ? sl[1] : g == 2 ? sl[2] : sl[3];
            int64_t msl = -1;
            double mj = 0.0, sj = 1.0;
            if (jg >= 0 && gl == 0) {
                msl = off_km[(int64_t)I * N + jg] + mirror[sg];
                mj = mu[jg];
                sj = sd[jg];
            }
            double acc[4];
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++) {
                double av = 0.0;
                if (lane < Lt && j[q4] >= 0) av = zt * (double)xs[(int64_t)j[q4] * ld + 4 * L4 + lane];
#pragma unroll
                for (int t = 0; t < T; t++)
#pragma unroll
                    for (int c = 0; c < 4; c++) av = __builtin_fma(zi[t][c], (double)xj[q4][t][c], av);
                acc[q4] = av;
            }
            // fold: after xor 32 lane half h5 holds entries {2h5, 2h5+1}; after xor
            // 16 quarter (h5, h4) holds entry 2h5 + h4; xor 8..1 finish the sum
            double v0 = (h5 ? acc[2] : acc[0]) + __shfl_xor(h5 ? acc[0] : acc[2], 32, 64);
            double v1 = (h5 ? acc[3] : acc[1]) + __shfl_xor(h5 ? acc[1] : acc[3], 32, 64);
            double w = (h4 ? v1 : v0) + __shfl_xor(h4 ? v0 : v1, 16, 64);
#pragma unroll
            for (int k = 8; k >= 1; k >>= 1) w += __shfl_xor(w, k, 64);
            if (jg >= 0 && gl == 0) {
                const float v = (float)((w - mj * s) / sj);
                V[2 * sg + 1] = v;
                V[2 * msl + 1] = v;
            }
        }
    };

    int nq = 0;  // wave-uniform
    for (int k0 = I + 1; k0 <= nb; k0 += 64) {
        const int k = k0 + lane;
        const int c = k <= nb ? orig_rm[i * nb1 + k] : 0;
        const int64_t off = c ? off_rm[i * nb1 + k] : 0;
        uint64_t m = __ballot(c > 0);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const int cnt = __shfl(c, src, 64);
            const int64_t o = __shfl(off, src, 64);
            if (nq + cnt > RQ) {
                flush(nq);
                nq = 0;
            }
            for (int t = lane; t < cnt; t += 64) q[nq + t] = o + t;
            nq += cnt;
        }
    }
    flush(nq);
}

// The list form holds row i in at most 8 pieces of four per lane (L <= 2048)
// and reads 16-byte aligned rows; other shapes refine the placed CSR instead.
bool refine_lists_supported(const float *xs, int64_t ld, int L)
{
    return L >= 1 && (L >> 2) <= 64 * 8 && ld % 4 == 0 && ((uintptr_t)xs & 15) == 0;
}

hipError_t launch_refine_lists(Entry *entries, const uint8_t *mirror, const int32_t *orig_rm, const int64_t *off_rm,
                               const int64_t *off_km, const float *xs, int64_t ld, const double *mu, const double *sd,
                               int N, int L, int nb, int row0, int row1, hipStream_t st)
{
    if (row0 < 0 || row1 > N || row0 > row1) return hipErrorInvalidValue;
    if (row0 == row1) return hipSuccess;
    if (!refine_lists_supported(xs, ld, L)) return hipErrorNotSupported;
    const int T = std::max(1, ((L >> 2) + 63) / 64);
    const dim3 grid((unsigned)(row1 - row0));
#define BNA_REFINE_LISTS(TT)                                                                       \
    case TT:                                                                                       \
        hipLaunchKernelGGL(refine_lists_kernel<TT>, grid, dim3(64), 0, st, entries, mirror, orig_rm, off_rm, off_km, \
                           xs, ld, mu, sd, N, L, nb, row0, row1);                                  \
        break;
    switch (T) {
        BNA_REFINE_LISTS(1) BNA_REFINE_LISTS(2) BNA_REFINE_LISTS(3) BNA_REFINE_LISTS(4)
        BNA_REFINE_LISTS(5) BNA_REFINE_LISTS(6) BNA_REFINE_LISTS(7) BNA_REFINE_LISTS(8)
    }
#undef BNA_REFINE_LISTS
    return hipGetLastError();
}

hipError_t launch_refine(const int32_t *R, const int32_t *C, float *V, const float *xs, int64_t ld,
                         const double *mu, const double *sd, int N, int L, int row0, int row1, int64_t nnz,
                         hipStream_t st)
{
    if (row0 < 0 || row1 > N || row0 >= row1) return row0 == row1 ? hipSuccess : hipErrorInvalidValue;
    const int64_t threads = (int64_t)(row1 - row0) * 64;
    const dim3 grid((unsigned)((threads + 255) / 256));
    // chunked form: one wave per chunk of the whole CSR; the waves of a launch
    // over fewer rows find their chunks gone and leave
    const int64_t cw = std::max<int64_t>(4, (nnz + REFINE_CHUNK - 1) / REFINE_CHUNK);
    const dim3 cgrid((unsigned)((cw + 3) / 4));
    // 16-byte loads need 16-byte aligned rows
    const bool vec = ld % 4 == 0 && ((uintptr_t)xs & 15) == 0;
    const int T = ((L >> 2) + 63) / 64;
    if (vec && T >= 1 && T <= 8) {
#define BNA_REFINE_ROWS(TT) \
    case TT: hipLaunchKernelGGL(refine_rows_kernel<TT>, cgrid, dim3(256), 0, st, R, C, V, xs, ld, mu, sd, N, L, row0, row1); break;
        switch (T) {
            BNA_REFINE_ROWS(1) BNA_REFINE_ROWS(2) BNA_REFINE_ROWS(3) BNA_REFINE_ROWS(4)
            BNA_REFINE_ROWS(5) BNA_REFINE_ROWS(6) BNA_REFINE_ROWS(7) BNA_REFINE_ROWS(8)
        }
#undef BNA_REFINE_ROWS
    } else if (vec)
        hipLaunchKernelGGL(refine_kernel<true>, grid, dim3(256), 0, st, R, C, V, xs, ld, mu, sd, N, L, row0, row1);
    else
        hipLaunchKernelGGL(refine_kernel<false>, grid, dim3(256), 0, st, R, C, V, xs, ld, mu, sd, N, L, row0, row1);
    return hipGetLastError();
}

hipError_t launch_recheck(const float *xs, int64_t ld, const double *mu, const double *sd, int N, int L, int nb,
                          const Candidate *cands,
                          const unsigned long long *cand_count, int64_t cand_cap, double thr,
                          int32_t *frag_live, Entry *entries, int64_t entry_cap, unsigned long long *stats,
                          int mode, double *r64_out, hipStream_t st)
{
    // enough waves to cover the device; candidates are grid-strided
    hipLaunchKernelGGL(recheck_kernel, dim3(2048), dim3(256), 0, st, xs, ld, mu, sd, N, L, nb, cands, cand_count,
                       cand_cap,
                       thr, frag_live, entries, entry_cap, stats, mode, r64_out);
    return hipGetLastError();
}

// Zeroes n 64-bit words.  A kernel rather than hipMemsetAsync: the build's
// launches may be captured into a HIP graph, where memset nodes are avoided.
__global__ __launch_bounds__(256) void zero_u64_kernel(unsigned long long *__restrict__ p, int64_t n)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) p[k] = 0ull;
}

hipError_t launch_zero_u64(void *p, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (unsigned long long *)p, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void narrow_rows_kernel(const int64_t *__restrict__ R64, int64_t n,
                                                           int32_t *__restrict__ R32)
{
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256)
        R32[k] = (int32_t)R64[k];
}

// R as int32 (the caller's CSR) from the scan's int64 offsets; nnz < 2^31 is
// checked by the caller.
hipError_t launch_narrow_rows(const int64_t *R64, int64_t n, int32_t *R32, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(narrow_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, st, R64, n, R32);
    return hipGetLastError();
}

hipError_t launch_frag_rows(const int32_t *orig_km, const int32_t *live_km, const int64_t *off_km, int N, int nb,
                            int32_t *orig_rm, int32_t *live_rm, int64_t *off_rm, int64_t *rowlen, hipStream_t st)
{
    hipError_t e = launch_zero_u64(rowlen, N, st);
    if (e != hipSuccess) return e;
    const int nb1 = nb + 1;
    hipLaunchKernelGGL(frag_rows_kernel, dim3((unsigned)((nb1 + TR - 1) / TR), (unsigned)((N + TR - 1) / TR)),
                       dim3(256), 0, st, orig_km, live_km, off_km, N, nb1, orig_rm, live_rm, off_rm,
                       (unsigned long long *)rowlen);
    return hipGetLastError();
}

hipError_t launch_scan(const int64_t *rowlen, int N, int64_t *R64, int32_t *R32, int64_t *bsum, hipStream_t st)
{
    const int nblocks = (N + SCAN_E - 1) / SCAN_E;
    hipLaunchKernelGGL(scan_sums_kernel, dim3(nblocks), dim3(SCAN_T), 0, st, rowlen, N, bsum);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(SCAN_T), 0, st, bsum, nblocks);
    hipLaunchKernelGGL(scan_down_kernel, dim3(nblocks), dim3(SCAN_T), 0, st, rowlen, N, bsum, nblocks, R64, R32);
    return hipGetLastError();
}

int scan_blocks(int N) { return (N + SCAN_E - 1) / SCAN_E; }

hipError_t launch_place(const int32_t *frag_orig, const int32_t *frag_live, const int64_t *frag_off,
                        const Entry *entries, const int64_t *R64, int N, int nb, int32_t *C, float *V,
                        hipStream_t st)
{
    const int64_t threads = (int64_t)N * 64;
    hipLaunchKernelGGL(place_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, frag_orig,
                       frag_live, frag_off, entries, R64, N, nb, C, V);
    return hipGetLastError();
}

}  // namespace bna
