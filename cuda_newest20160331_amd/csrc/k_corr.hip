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
// global-address-space views, so panel loads compile to global_load_dwordx4
typedef __attribute__((address_space(1))) const unsigned char gbyte;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 guint4;

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

// ---------------------------------------------------------------- epilogue
// One 32x32 accumulator block at a time.  The block indices are template
// arguments so every accumulator access is constant-indexed and the 256
// accumulators stay in registers.  Element e of a block sits at tile row
// wr0 + 32 RB + 8 (e / 4) + 4 half + e % 4 and tile column wc0 + 32 CB + lane % 32.
struct EpiCtx {
    int N, gi0, gj0, wr0, wc0, half, lrow;
    bool diag;
    float keep, sure;
};

struct EmitOut {
    Entry *entries;
    Candidate *cands;
    unsigned long long *cand_count;
    int64_t cand_cap;
    int64_t base;
    int total;
    bool fits;
};

__device__ __forceinline__ bool epi_in(const EpiCtx &c, int lr, int lc) {
    return (c.gi0 + lr < c.N) && (c.gj0 + lc < c.N) && (!c.diag || lr < lc);
}

template <int RB, int CB>
__device__ __forceinline__ void epi_hist(const f32x16_t &a, const EpiCtx &c, uint32_t *hist) {
    const int lc = c.wc0 + CB * 32 + c.lrow;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        const float v = a[e];
        if (epi_in(c, lr, lc) && v == v) {
            int bin = (int)floorf((v + 1.0f) * (HIST_BINS / 2));
            bin = bin < 0 ? 0 : (bin >= HIST_BINS ? HIST_BINS - 1 : bin);
            atomicAdd(&hist[bin], 1u);
        }
    }
}

template <int RB, int CB>
__device__ __forceinline__ void epi_mark(const f32x16_t &a, const EpiCtx &c, uint32_t *rowbits, uint32_t *colbits) {
    const int lc = c.wc0 + CB * 32 + c.lrow;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        if (a[e] > c.keep && epi_in(c, lr, lc)) {
            atomicOr(&rowbits[lr * 8 + (lc >> 5)], 1u << (lc & 31));
            atomicOr(&colbits[lc * 8 + (lr >> 5)], 1u << (lr & 31));
        }
    }
}

template <int RB, int CB>
__device__ __forceinline__ void epi_emit(const f32x16_t &a, const EpiCtx &c, const uint32_t *rowbits,
                                         const uint32_t *colbits, const int *rowpre, const int *colpre,
                                         const EmitOut &o) {
    const int lc = c.wc0 + CB * 32 + c.lrow;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        const float v = a[e];
        if (v > c.keep && epi_in(c, lr, lc)) {
            const int rw = lr * 8 + (lc >> 5), cw = lc * 8 + (lr >> 5);
            const int64_t p = o.base + rowpre[rw] + __popc(rowbits[rw] & ((1u << (lc & 31)) - 1u));
            const int64_t q = o.base + o.total + colpre[cw] + __popc(colbits[cw] & ((1u << (lr & 31)) - 1u));
            if (o.fits) {
                o.entries[p] = Entry{c.gj0 + lc, v};
                o.entries[q] = Entry{c.gi0 + lr, v};
            }
            if (!(v > c.sure)) {
                const unsigned long long ci = atomicAdd(o.cand_count, 1ull);
                if ((int64_t)ci < o.cand_cap) o.cands[ci] = Candidate{c.gi0 + lr, c.gj0 + lc, p, q, v, 0};
            }
        }
    }
}

#define BNA_FOR_BLOCKS(F, ...)                                                                     \
    do {                                                                                           \
        F<0, 0>(acc[0][0], __VA_ARGS__); F<0, 1>(acc[0][1], __VA_ARGS__);                          \
        F<0, 2>(acc[0][2], __VA_ARGS__); F<0, 3>(acc[0][3], __VA_ARGS__);                          \
        F<1, 0>(acc[1][0], __VA_ARGS__); F<1, 1>(acc[1][1], __VA_ARGS__);                          \
        F<1, 2>(acc[1][2], __VA_ARGS__); F<1, 3>(acc[1][3], __VA_ARGS__);                          \
        F<2, 0>(acc[2][0], __VA_ARGS__); F<2, 1>(acc[2][1], __VA_ARGS__);                          \
        F<2, 2>(acc[2][2], __VA_ARGS__); F<2, 3>(acc[2][3], __VA_ARGS__);                          \
        F<3, 0>(acc[3][0], __VA_ARGS__); F<3, 1>(acc[3][1], __VA_ARGS__);                          \
        F<3, 2>(acc[3][2], __VA_ARGS__); F<3, 3>(acc[3][3], __VA_ARGS__);                          \
    } while (0)

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

    // Global sources, one per (side, array): uniform per workgroup.  Each
    // thread owns NPRE 16-byte chunks per stage at fixed offsets within the
    // panels; the stage index only adds kt * KB elements.
    const int64_t row_stride = (int64_t)L_pad * C::ESZ;
    const gbyte *gsrc[2 * C::NARR];
    gsrc[0] = (const gbyte *)sd.a0 + (int64_t)I * TILE * row_stride;
    gsrc[C::NARR] = (const gbyte *)sd.a0 + (int64_t)J * TILE * row_stride;
    if constexpr (C::NARR == 2) {
        gsrc[1] = (const gbyte *)sd.a1 + (int64_t)I * TILE * row_stride;
        gsrc[3] = (const gbyte *)sd.a1 + (int64_t)J * TILE * row_stride;
    }
    uint32_t goff[C::CH];
    uint32_t loff[C::CH];
#pragma unroll
    for (int m = 0; m < C::CH; m++) {
        const int id = m * CT_THREADS + tid;
        const int row = id / C::CH, c = id % C::CH;
        goff[m] = (uint32_t)(row * row_stride + c * 16);
        loff[m] = (uint32_t)lds_off<PREC>(0, 0, row, c);
    }

    u32x4 pre[C::NPRE];
#define BNA_GLOAD(KT)                                                                              \
    do {                                                                                           \
        const uint32_t kofs_ = (uint32_t)(KT) * (KB * C::ESZ);                                     \
        _Pragma("unroll") for (int sa = 0; sa < 2 * C::NARR; sa++)                                 \
            _Pragma("unroll") for (int m = 0; m < C::CH; m++)                                      \
                pre[sa * C::CH + m] = *(const guint4 *)(gsrc[sa] + goff[m] + kofs_);               \
    } while (0)
#define BNA_LSTORE(STG_)                                                                          \
    do {                                                                                           \
        unsigned char *b_ = smem + (STG_) * C::STAGE;                                             \
        _Pragma("unroll") for (int sa = 0; sa < 2 * C::NARR; sa++)                                 \
            _Pragma("unroll") for (int m = 0; m < C::CH; m++)                                      \
                *(u32x4 *)(b_ + sa * C::PANEL + loff[m]) = pre[sa * C::CH + m];                    \
    } while (0)

    f32x16_t acc[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[rb][cb][e] = 0.0f;

    BNA_GLOAD(0);
    BNA_LSTORE(0);
    __syncthreads();

    const int lrow = lane & 31;
    const int half = lane >> 5;

    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        // the last stage reloads itself so the loop body has no branch; the
        // copy lands in the idle buffer and is never read
        BNA_GLOAD(kt + 1 < nk ? kt + 1 : kt);
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
        BNA_LSTORE(cur ^ 1);
        __syncthreads();
    }
#undef BNA_GLOAD
#undef BNA_LSTORE

    // ------------------------------------------------------------ epilogue
    EpiCtx ec;
    ec.N = sd.N;
    ec.gi0 = I * TILE;
    ec.gj0 = J * TILE;
    ec.wr0 = wr0;
    ec.wc0 = wc0;
    ec.half = half;
    ec.lrow = lrow;
    ec.diag = (I == J);
    ec.keep = sd.keep_thr;
    ec.sure = sd.sure_thr;

    if constexpr (HIST) {
        uint32_t *hist = (uint32_t *)smem;
        for (int k = tid; k < HIST_BINS; k += CT_THREADS) hist[k] = 0;
        __syncthreads();
        BNA_FOR_BLOCKS(epi_hist, ec, hist);
        __syncthreads();
        for (int k = tid; k < HIST_BINS; k += CT_THREADS)
            if (hist[k]) atomicAdd(&sd.hist[k], (unsigned long long)hist[k]);
        return;
    } else {
        uint32_t *rowbits = (uint32_t *)smem;        // [256][8] tile-row bitmaps
        uint32_t *colbits = rowbits + TILE * 8;      // [256][8] tile-column bitmaps
        int *rowpre = (int *)(colbits + TILE * 8);   // [256][8] slot of each bitmap word
        int *colpre = rowpre + TILE * 8;             // [256][8]
        int *wsum = colpre + TILE * 8;               // [4]
        int *wsum2 = wsum + 4;                       // [4]
        unsigned long long *sbase = (unsigned long long *)(wsum2 + 4);
        for (int k = tid; k < TILE * 8 * 2; k += CT_THREADS) rowbits[k] = 0;
        __syncthreads();

        BNA_FOR_BLOCKS(epi_mark, ec, rowbits, colbits);
        __syncthreads();

        int rw[8], cw[8];
        int rc = 0, cc = 0;
#pragma unroll
        for (int w = 0; w < 8; w++) {
            rw[w] = __popc(rowbits[tid * 8 + w]);
            cw[w] = __popc(colbits[tid * 8 + w]);
            rc += rw[w];
            cc += cw[w];
        }
        int total = 0, total2 = 0;
        const int ro = block_excl_scan(rc, wsum, total);
        const int co = block_excl_scan(cc, wsum2, total2);
        {
            int a = ro, b2 = co;
#pragma unroll
            for (int w = 0; w < 8; w++) {
                rowpre[tid * 8 + w] = a;
                colpre[tid * 8 + w] = b2;
                a += rw[w];
                b2 += cw[w];
            }
        }
        if (tid == 0) *sbase = atomicAdd(sd.entry_count, (unsigned long long)(2 * total));
        __syncthreads();
        const int64_t base = (int64_t)*sbase;
        const bool fits = base + 2 * (int64_t)total <= sd.entry_cap;

        const int nb1 = sd.nb + 1;
        if (ec.gi0 + tid < ec.N) {
            const int64_t f = (int64_t)(ec.gi0 + tid) * nb1 + (J + 1);
            sd.frag_orig[f] = rc;
            sd.frag_live[f] = rc;
            sd.frag_off[f] = base + ro;
        }
        if (ec.gj0 + tid < ec.N) {
            const int64_t f = (int64_t)(ec.gj0 + tid) * nb1 + I;
            sd.frag_orig[f] = cc;
            sd.frag_live[f] = cc;
            sd.frag_off[f] = base + total + co;
        }
        if (total == 0) return;
        EmitOut eo{sd.entries, sd.cands, sd.cand_count, sd.cand_cap, base, total, fits};
        BNA_FOR_BLOCKS(epi_emit, ec, rowbits, colbits, rowpre, colpre, eo);
    }
}

template <int PREC, int HIST>
static hipError_t launch_one(const SubjDesc *subj_dev, const TileItem *items, int64_t n_items, hipStream_t st)
{
    using C = CorrCfg<PREC>;
    constexpr size_t epi = (size_t)TILE * 8 * 4 * 4 + 64;
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
