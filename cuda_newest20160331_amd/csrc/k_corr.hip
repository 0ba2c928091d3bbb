// k_corr.hip -- kernel (b) and the first half of (c): blocked Pearson correlation
// as an MFMA tile product, fused with the r threshold and with the compaction of
// surviving pairs.  The dense N x N matrix is never written.
//
// Replaces the reference's per-tile cublasSgemm + D2H + host threshold + host
// CSR scan (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:87-99, CorMat.cpp:683-703,
// 889-918).  Only upper-triangle tiles (I <= J) run; each surviving pair (i, j)
// is written twice, once in tile-row order (for row i's neighbours j > i) and
// once in tile-column order (for row j's neighbours i < j), so the CSR rows can
// be assembled later without a sort.
//
// Work split: a 256 x 256 tile per workgroup of 4 waves; each wave owns a
// 128 x 128 sub-tile held as 4 x 4 MFMA 32x32 accumulator blocks.  K runs in
// stages of 32 time points, double-buffered in LDS, with the next stage's global
// loads in flight in registers while the current stage feeds the MFMAs.
//
// Precisions (template PREC):
//   0 FP32    v_mfma_f32_32x32x2_f32,  z32 operands
//   1 BF16    v_mfma_f32_32x32x16_bf16, hi operands
//   2 BF16X3  v_mfma_f32_32x32x16_bf16, hi*hi + hi*lo + lo*hi
// All accumulate in FP32.
//
// Epilogue: each kept pair (r > keep_thr) sets a bit in a tile-row and a
// tile-column bitmap in LDS; per-row and per-column popcounts, two block scans
// and one global atomic give every pair its slot in both lists.  Pairs with
// r <= sure_thr are also queued as candidates for the FP64 recheck.
#include "common.hpp"

namespace bna {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int CT_THREADS = 256;
constexpr int KB = 32;  // time points per stage

template <int PREC>
struct CorrCfg {
    static constexpr int NARR = (PREC == 2) ? 2 : 1;       // arrays per panel
    static constexpr int ROWB = (PREC == 0) ? 128 : 64;    // bytes per panel row per stage
    static constexpr int CH = ROWB / 16;                   // 16-byte chunks per row
    static constexpr int ESZ = (PREC == 0) ? 4 : 2;        // element size
    static constexpr int PANEL = TILE * ROWB;              // bytes per array per side
    static constexpr int STAGE = 2 * NARR * PANEL;         // A + B
    static constexpr int NPRE = 2 * NARR * CH;             // uint4 per thread per stage
};

// XOR swizzle of the 16-byte chunk index so that the 16 lanes of each
// ds_read_b128 phase touch 16 different bank groups.
template <int CH>
__device__ __forceinline__ int swz(int row) {
    return (CH == 4) ? ((row >> 2) & 3) : ((row >> 1) & 7);
}

template <int PREC>
__device__ __forceinline__ int lds_off(int side, int arr, int row, int c) {
    using C = CorrCfg<PREC>;
    return ((side * C::NARR + arr) * TILE + row) * C::ROWB + ((c ^ swz<C::CH>(row)) << 4);
}

// Block-wide exclusive scan of one int per thread (256 threads).
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int pre = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < CT_THREADS / 64; k++) {
        int s = wsum[k];
        if (k < w) pre += s;
        total += s;
    }
    return pre + inc - v;
}

__device__ __forceinline__ int popc_below(const uint32_t *bits, int pos) {
    int n = 0;
    const int w = pos >> 5;
    for (int k = 0; k < w; k++) n += __popc(bits[k]);
    n += __popc(bits[w] & ((1u << (pos & 31)) - 1u));
    return n;
}

template <int PREC, int HIST>
__global__ __launch_bounds__(CT_THREADS, 1) void corr_tile_kernel(const SubjDesc *__restrict__ subjs,
                                                                  const TileItem *__restrict__ items,
                                                                  int64_t n_items)
{
    using C = CorrCfg<PREC>;
    extern __shared__ __align__(16) unsigned char smem[];

    // XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a contiguous
    // run of the (band-ordered) tile table so tiles that share panels share an L2.
    const int64_t per_xcd = (n_items + 7) / 8;
    const int64_t b = blockIdx.x;
    const int64_t logical = (b & 7) * per_xcd + (b >> 3);
    if (logical >= n_items) return;
    const TileItem it = items[logical];
    const SubjDesc sd = subjs[it.subj];
    const int I = it.I, J = it.J;
    const int L_pad = sd.L_pad;
    const int nk = L_pad / KB;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr0 = (wave >> 1) * 128;
    const int wc0 = (wave & 1) * 128;

    // global sources: [side][arr]
    const unsigned char *src[2][2];
    src[0][0] = (const unsigned char *)sd.a0 + (int64_t)I * TILE * L_pad * C::ESZ;
    src[1][0] = (const unsigned char *)sd.a0 + (int64_t)J * TILE * L_pad * C::ESZ;
    src[0][1] = (C::NARR == 2) ? (const unsigned char *)sd.a1 + (int64_t)I * TILE * L_pad * C::ESZ : nullptr;
    src[1][1] = (C::NARR == 2) ? (const unsigned char *)sd.a1 + (int64_t)J * TILE * L_pad * C::ESZ : nullptr;
    const int64_t row_stride = (int64_t)L_pad * C::ESZ;

    uint4 pre[C::NPRE];
    auto gload = [&](int kt) {
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int a = 0; a < C::NARR; a++)
#pragma unroll
                for (int m = 0; m < C::CH; m++) {
                    const int id = m * CT_THREADS + tid;
                    const int row = id / C::CH, c = id % C::CH;
                    pre[(s * C::NARR + a) * C::CH + m] =
                        *(const uint4 *)(src[s][a] + row * row_stride + (int64_t)kt * KB * C::ESZ + c * 16);
                }
    };
    auto lstore = [&](int stage) {
        unsigned char *base = smem + stage * C::STAGE;
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int a = 0; a < C::NARR; a++)
#pragma unroll
                for (int m = 0; m < C::CH; m++) {
                    const int id = m * CT_THREADS + tid;
                    const int row = id / C::CH, c = id % C::CH;
                    *(uint4 *)(base + lds_off<PREC>(s, a, row, c)) = pre[(s * C::NARR + a) * C::CH + m];
                }
    };

    f32x16_t acc[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[rb][cb][e] = 0.0f;

    gload(0);
    lstore(0);
    __syncthreads();

    const int lrow = lane & 31;
    const int half = lane >> 5;

    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload(kt + 1);
        const unsigned char *base = smem + cur * C::STAGE;
        if constexpr (PREC == 0) {
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
                float4 af[4], bfv[4];
                const int c = half * 4 + cc;  // lane half h covers k = 16h .. 16h+15
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
                    af[rb] = *(const float4 *)(base + lds_off<PREC>(0, 0, wr0 + rb * 32 + lrow, c));
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
                    bfv[cb] = *(const float4 *)(base + lds_off<PREC>(1, 0, wc0 + cb * 32 + lrow, c));
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int rb = 0; rb < 4; rb++)
#pragma unroll
                        for (int cb = 0; cb < 4; cb++)
                            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[rb][q], bfv[cb][q], acc[rb][cb], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const int c = half * 2 + kk;  // lane half h covers k = 16h .. 16h+15
                bf16x8_t ah[4], bh[4], al[4], bl[4];
#pragma unroll
                for (int rb = 0; rb < 4; rb++) {
                    ah[rb] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(base + lds_off<PREC>(0, 0, wr0 + rb * 32 + lrow, c)));
                    if constexpr (PREC == 2)
                        al[rb] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(base + lds_off<PREC>(0, 1, wr0 + rb * 32 + lrow, c)));
                }
#pragma unroll
                for (int cb = 0; cb < 4; cb++) {
                    bh[cb] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(base + lds_off<PREC>(1, 0, wc0 + cb * 32 + lrow, c)));
                    if constexpr (PREC == 2)
                        bl[cb] = __builtin_bit_cast(bf16x8_t, *(const uint4 *)(base + lds_off<PREC>(1, 1, wc0 + cb * 32 + lrow, c)));
                }
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                if constexpr (PREC == 2) {
#pragma unroll
                    for (int rb = 0; rb < 4; rb++)
#pragma unroll
                        for (int cb = 0; cb < 4; cb++)
                            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[rb], bl[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
                    for (int rb = 0; rb < 4; rb++)
#pragma unroll
                        for (int cb = 0; cb < 4; cb++)
                            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[rb], bh[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        }
        if (kt + 1 < nk) lstore(cur ^ 1);
        __syncthreads();
    }

    // ------------------------------------------------------------ epilogue
    const int N = sd.N;
    const int gi0 = I * TILE, gj0 = J * TILE;
    const bool diag = (I == J);

    if constexpr (HIST) {
        uint32_t *hist = (uint32_t *)smem;
        for (int k = tid; k < HIST_BINS; k += CT_THREADS) hist[k] = 0;
        __syncthreads();
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int lr = wr0 + rb * 32 + (e >> 2) * 8 + half * 4 + (e & 3);
                    const int lc = wc0 + cb * 32 + lrow;
                    const float v = acc[rb][cb][e];
                    const bool inb = (gi0 + lr < N) && (gj0 + lc < N) && (!diag || lr < lc) && (v == v);
                    if (inb) {
                        int bin = (int)floorf((v + 1.0f) * (HIST_BINS / 2));
                        bin = bin < 0 ? 0 : (bin >= HIST_BINS ? HIST_BINS - 1 : bin);
                        atomicAdd(&hist[bin], 1u);
                    }
                }
        __syncthreads();
        for (int k = tid; k < HIST_BINS; k += CT_THREADS)
            if (hist[k]) atomicAdd(&sd.hist[k], (unsigned long long)hist[k]);
        return;
    } else {
        uint32_t *rowbits = (uint32_t *)smem;        // [256][8]
        uint32_t *colbits = rowbits + TILE * 8;      // [256][8]
        int *rowoff = (int *)(colbits + TILE * 8);   // [256]
        int *coloff = rowoff + TILE;                 // [256]
        int *wsum = coloff + TILE;                   // [4]
        int *wsum2 = wsum + 4;                       // [4]
        unsigned long long *sbase = (unsigned long long *)(wsum2 + 4);
        for (int k = tid; k < TILE * 8 * 2; k += CT_THREADS) rowbits[k] = 0;
        __syncthreads();

        const float keep = sd.keep_thr;
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int lr = wr0 + rb * 32 + (e >> 2) * 8 + half * 4 + (e & 3);
                    const int lc = wc0 + cb * 32 + lrow;
                    const float v = acc[rb][cb][e];
                    const bool kp = (v > keep) && (gi0 + lr < N) && (gj0 + lc < N) && (!diag || lr < lc);
                    if (kp) {
                        atomicOr(&rowbits[lr * 8 + (lc >> 5)], 1u << (lc & 31));
                        atomicOr(&colbits[lc * 8 + (lr >> 5)], 1u << (lr & 31));
                    }
                }
        __syncthreads();

        int rc = 0, cc = 0;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            rc += __popc(rowbits[tid * 8 + w]);
            cc += __popc(colbits[tid * 8 + w]);
        }
        int total = 0, total2 = 0;
        const int ro = block_excl_scan(rc, wsum, total);
        const int co = block_excl_scan(cc, wsum2, total2);
        rowoff[tid] = ro;
        coloff[tid] = co;
        if (tid == 0) *sbase = atomicAdd(sd.entry_count, (unsigned long long)(2 * total));
        __syncthreads();
        const int64_t base = (int64_t)*sbase;
        const bool fits = base + 2 * (int64_t)total <= sd.entry_cap;

        const int nb1 = sd.nb + 1;
        if (gi0 + tid < N) {
            const int64_t f = (int64_t)(gi0 + tid) * nb1 + (J + 1);
            sd.frag_orig[f] = rc;
            sd.frag_live[f] = rc;
            sd.frag_off[f] = base + ro;
        }
        if (gj0 + tid < N) {
            const int64_t f = (int64_t)(gj0 + tid) * nb1 + I;
            sd.frag_orig[f] = cc;
            sd.frag_live[f] = cc;
            sd.frag_off[f] = base + total + co;
        }
        if (total == 0) return;

        const float sure = sd.sure_thr;
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int e = 0; e < 16; e++) {
                    const int lr = wr0 + rb * 32 + (e >> 2) * 8 + half * 4 + (e & 3);
                    const int lc = wc0 + cb * 32 + lrow;
                    const float v = acc[rb][cb][e];
                    const bool kp = (v > keep) && (gi0 + lr < N) && (gj0 + lc < N) && (!diag || lr < lc);
                    if (kp) {
                        const int64_t p = base + rowoff[lr] + popc_below(rowbits + lr * 8, lc);
                        const int64_t q = base + total + coloff[lc] + popc_below(colbits + lc * 8, lr);
                        if (fits) {
                            sd.entries[p] = Entry{gj0 + lc, v};
                            sd.entries[q] = Entry{gi0 + lr, v};
                        }
                        if (!(v > sure)) {
                            const unsigned long long ci = atomicAdd(sd.cand_count, 1ull);
                            if ((int64_t)ci < sd.cand_cap)
                                sd.cands[ci] = Candidate{gi0 + lr, gj0 + lc, p, q, v, 0};
                        }
                    }
                }
    }
}

template <int PREC, int HIST>
static hipError_t launch_one(const SubjDesc *subj_dev, const TileItem *items, int64_t n_items, hipStream_t st)
{
    using C = CorrCfg<PREC>;
    constexpr size_t epi = (size_t)TILE * 8 * 4 * 2 + TILE * 4 * 2 + 64;
    constexpr size_t hst = (size_t)HIST_BINS * 4;
    constexpr size_t need = 2 * (size_t)C::STAGE;
    constexpr size_t shm = need > epi ? (need > hst ? need : hst) : (epi > hst ? epi : hst);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)corr_tile_kernel<PREC, HIST>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int64_t per_xcd = (n_items + 7) / 8;
    const int64_t grid = per_xcd * 8;
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((corr_tile_kernel<PREC, HIST>), dim3((unsigned)grid), dim3(CT_THREADS), shm, st, subj_dev,
                       items, n_items);
    return hipGetLastError();
}

hipError_t launch_corr(int prec, int hist_mode, const SubjDesc *subj_dev, const TileItem *items,
                       int64_t n_items, hipStream_t st)
{
    switch (prec * 2 + (hist_mode ? 1 : 0)) {
    case 0: return launch_one<0, 0>(subj_dev, items, n_items, st);
    case 1: return launch_one<0, 1>(subj_dev, items, n_items, st);
    case 2: return launch_one<1, 0>(subj_dev, items, n_items, st);
    case 3: return launch_one<1, 1>(subj_dev, items, n_items, st);
    case 4: return launch_one<2, 0>(subj_dev, items, n_items, st);
    case 5: return launch_one<2, 1>(subj_dev, items, n_items, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace bna
