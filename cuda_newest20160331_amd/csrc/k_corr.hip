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
// 128 x 128 sub-tile held as 4 x 4 MFMA 32x32 accumulator blocks (256 AGPRs).
//
// Operand panels are k-blocked: for each block of 16 time points kb and each
// voxel row r there is one UNIT-byte record at ((kb * N_pad) + r) * UNIT, so
// one stage of a 256-row panel is a contiguous 256 * UNIT bytes.  Stages move
// global -> LDS with global_load_lds_dwordx4 (no registers in between) into a
// ring of 4 slots; the load for stage kt + 4 is issued while stage kt is
// multiplied, so every load has three stages of MFMA work to arrive in.  The
// MFMA fragments are double-buffered in registers: the LDS reads of stage
// kt + 1 are issued before the MFMAs of stage kt.
//
// Precisions (template PREC), all accumulating in FP32:
//   0 FP32    v_mfma_f32_32x32x2_f32,   UNIT 64: z[16] fp32
//   1 BF16    v_mfma_f32_32x32x16_bf16, UNIT 32: hi[16]
//   2 BF16X3  v_mfma_f32_32x32x16_bf16, UNIT 64: hi[16] lo[16]; hi*hi + hi*lo + lo*hi
//
// Epilogue: each kept pair (r > keep_thr) sets a bit in a tile-row and a
// tile-column bitmap in LDS; per-row and per-column popcounts, two block scans
// and one global atomic give every pair its slot in both lists.  Pairs with
// r <= sure_thr are also queued as candidates for the FP64 recheck.
#include "common.hpp"

#include <atomic>

namespace bna {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) void gvoid;
typedef __attribute__((address_space(3))) void lvoid;

constexpr int CT_THREADS = 256;
constexpr int NSLOT = 4;  // LDS ring of k16 stages
// Epilogue LDS: bitmaps (16 KB), offsets (2 KB), scan partials, then the lists.
constexpr int EPI_LIST_OFF = 20 * 1024;

template <int PREC>
struct CorrCfg {
    static constexpr int UNIT = (PREC == 1) ? 32 : 64;   // bytes per (k16 block, row)
    static constexpr int CH = UNIT / 16;                 // 16-byte chunks per row
    static constexpr int PANEL = TILE * UNIT;            // bytes of one side of a stage
    static constexpr int SLOT = 2 * PANEL;               // A + B
    static constexpr int DMA_PER_SIDE = PANEL / 1024 / 4;  // 1 KB wave instructions per wave per side
};

// XOR swizzle of the 16-byte chunk index within a row so that the 16 lanes of
// each ds_read_b128 phase (16 consecutive rows, one chunk each) cover 16
// different bank groups.
template <int CH>
__device__ __forceinline__ int swz(int row) {
    return (CH == 4) ? ((row >> 2) & 3) : ((row >> 3) & 1);
}

// s_waitcnt vmcnt <= n and lgkmcnt == 0 (gfx9 encoding: vmcnt bits 3:0 and
// 15:14, expcnt 6:4 left unconstrained, lgkmcnt 11:8).  The lgkmcnt part makes
// sure this wave's LDS reads of a slot have completed before the barrier that
// lets the slot be refilled.
#define BNA_WAIT_VMCNT(n) \
    __builtin_amdgcn_s_waitcnt((((n) & 0xf) | (((n) >> 4) << 14)) | (0x7 << 4))

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

// ---------------------------------------------------------------- epilogue
// One 32x32 accumulator block at a time.  The block indices are template
// arguments so every accumulator access is constant-indexed and the 256
// accumulators stay in registers.  Element e of a block sits at tile row
// wr0 + 32 RB + 8 (e / 4) + 4 half + e % 4 and tile column wc0 + 32 CB + lane % 32.
struct EpiCtx {
    int N, gi0, gj0, wr0, wc0, half, lrow;
    bool diag;
    float keep;
};

// Per-wave kept-pair list: the first LIST_LDS entries in LDS, the rest (dense
// tiles only) in a global overflow area reserved on first need.
constexpr int LIST_LDS = 2048;
constexpr int WAVE_MAX = 128 * 128;  // a wave's whole sub-tile

// Records are stored packed in one 64-bit word (tile row | tile column << 16
// in the low half, r bits in the high half), with explicit LDS and global
// address spaces.
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(1))) uint64_t glb_u64;
struct WaveList {
    lds_u64 *lds;
    glb_u64 *ovf;    // null until reserved
    int n;           // entries so far (wave-uniform)
    bool lost;       // the overflow pool was exhausted: the host enlarges it and reruns
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

// The single pass over the accumulators: marks the block's kept pairs in both
// bitmaps and appends them to the wave's list at ballot-compacted positions.
template <int RB, int CB>
__device__ __forceinline__ void epi_mark(const f32x16_t &a, const EpiCtx &c, uint32_t *rowbits, uint32_t *colbits,
                                         WaveList &wl, KeptPair *ovf_pool, unsigned long long *ovf_count,
                                         int64_t ovf_cap)
{
    if (wl.ovf == nullptr && !wl.lost && wl.n + 1024 > LIST_LDS) {
        // this block could pass the LDS capacity: reserve the wave's worst case
        unsigned long long base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(ovf_count, (unsigned long long)(WAVE_MAX - LIST_LDS));
        base = __shfl(base, 0, 64);
        if ((int64_t)(base + (WAVE_MAX - LIST_LDS)) <= ovf_cap)
            wl.ovf = (glb_u64 *)(ovf_pool + base);
        else
            wl.lost = true;
    }
    const int lc = c.wc0 + CB * 32 + c.lrow;
    const uint64_t lt = (1ull << (threadIdx.x & 63)) - 1ull;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        const float v = a[e];
        const bool kp = v > c.keep && epi_in(c, lr, lc);
        const uint64_t m = __ballot(kp);
        if (kp) {
            atomicOr(&rowbits[lr * 8 + (lc >> 5)], 1u << (lc & 31));
            atomicOr(&colbits[lc * 8 + (lr >> 5)], 1u << (lr & 31));
            const int k = wl.n + __popcll(m & lt);
            const uint64_t rec = (uint64_t)((uint32_t)lr | ((uint32_t)lc << 16)) |
                                 ((uint64_t)__builtin_bit_cast(uint32_t, v) << 32);
            if (k < LIST_LDS) wl.lds[k] = rec;
            else if (wl.ovf) wl.ovf[k - LIST_LDS] = rec;
        }
        wl.n += __popcll(m);
    }
}

// Same marking for a tile that lies inside the matrix and off the diagonal,
// where every element is a pair to test.  The sixteen comparisons become
// sixteen wave masks first; a block none of whose elements is kept (most of
// them at useful thresholds) costs those compares and one OR, and inside a
// block only the elements some lane keeps take the recording branch.
template <int RB, int CB>
__device__ __forceinline__ void epi_mark_inner(const f32x16_t &a, const EpiCtx &c, uint32_t *rowbits,
                                               uint32_t *colbits, WaveList &wl, KeptPair *ovf_pool,
                                               unsigned long long *ovf_count, int64_t ovf_cap)
{
    uint64_t m[16];
    uint64_t any = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        m[e] = __ballot(a[e] > c.keep);
        any |= m[e];
    }
    if (any == 0) return;
    if (wl.ovf == nullptr && !wl.lost && wl.n + 1024 > LIST_LDS) {
        unsigned long long base = 0;
        if ((threadIdx.x & 63) == 0) base = atomicAdd(ovf_count, (unsigned long long)(WAVE_MAX - LIST_LDS));
        base = __shfl(base, 0, 64);
        if ((int64_t)(base + (WAVE_MAX - LIST_LDS)) <= ovf_cap)
            wl.ovf = (glb_u64 *)(ovf_pool + base);
        else
            wl.lost = true;
    }
    const int lc = c.wc0 + CB * 32 + c.lrow;
    const uint64_t lt = (1ull << (threadIdx.x & 63)) - 1ull;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        if (m[e] == 0) continue;
        const float v = a[e];
        if (v > c.keep) {
            const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
            atomicOr(&rowbits[lr * 8 + (lc >> 5)], 1u << (lc & 31));
            atomicOr(&colbits[lc * 8 + (lr >> 5)], 1u << (lr & 31));
            const int k = wl.n + __popcll(m[e] & lt);
            const uint64_t rec = (uint64_t)((uint32_t)lr | ((uint32_t)lc << 16)) |
                                 ((uint64_t)__builtin_bit_cast(uint32_t, v) << 32);
            if (k < LIST_LDS) wl.lds[k] = rec;
            else if (wl.ovf) wl.ovf[k - LIST_LDS] = rec;
        }
        wl.n += __popcll(m[e]);
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
    extern __shared__ __align__(1024) unsigned char smem[];

    // XCD-aware order: workgroup b runs on XCD b % 8; give each XCD a contiguous
    // run of the (band-ordered) tile table so tiles that share panels share an L2.
    const int64_t per_xcd = (n_items + 7) / 8;
    const int64_t b = blockIdx.x;
    const int64_t logical = (b & 7) * per_xcd + (b >> 3);
    if (logical >= n_items) return;
    const TileItem it = items[logical];
    const SubjDesc sd = subjs[it.subj];
    const int I = it.I, J = it.J;
    const int N_pad = sd.N_pad;
    const int nk = sd.L_pad / 16;  // k16 stages

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr0 = (wave >> 1) * 128;
    const int wc0 = (wave & 1) * 128;
    const int lrow = lane & 31;
    const int half = lane >> 5;

    // DMA addressing.  A wave instruction fills one contiguous KB of LDS, lane l
    // the 16 bytes at l * 16: row l / CH of the KB's rows, physical chunk l % CH,
    // which holds logical chunk (l % CH) ^ swz(row).  Wave w fills rows
    // [w * 64, w * 64 + 64) of each side.
    const unsigned char *gA = (const unsigned char *)sd.a0 + (int64_t)I * TILE * C::UNIT;
    const unsigned char *gB = (const unsigned char *)sd.a0 + (int64_t)J * TILE * C::UNIT;
    const int64_t kstride = (int64_t)N_pad * C::UNIT;  // bytes between k16 blocks
    constexpr int ROWS_PER_KB = 1024 / C::UNIT;
    uint32_t dsrc[C::DMA_PER_SIDE];  // byte offset within a stage panel, per instruction
#pragma unroll
    for (int q = 0; q < C::DMA_PER_SIDE; q++) {
        const int row = wave * (TILE / 4) + q * ROWS_PER_KB + lane / C::CH;
        const int pc = lane % C::CH;
        dsrc[q] = (uint32_t)(row * C::UNIT + ((pc ^ swz<C::CH>(row)) << 4));
    }
#define BNA_DMA(STAGE_IDX)                                                                         \
    do {                                                                                           \
        const int st_ = (STAGE_IDX) < nk ? (STAGE_IDX) : nk - 1;                                   \
        unsigned char *slot_ = smem + ((STAGE_IDX) % NSLOT) * C::SLOT;                             \
        const unsigned char *a_ = gA + st_ * kstride;                                              \
        const unsigned char *b_ = gB + st_ * kstride;                                              \
        _Pragma("unroll") for (int q = 0; q < C::DMA_PER_SIDE; q++) {                              \
            unsigned char *dA = slot_ + (wave * C::DMA_PER_SIDE + q) * 1024;                       \
            __builtin_amdgcn_global_load_lds((gvoid *)(a_ + dsrc[q]), (lvoid *)dA, 16, 0, 0);      \
            __builtin_amdgcn_global_load_lds((gvoid *)(b_ + dsrc[q]), (lvoid *)(dA + C::PANEL), 16, 0, 0); \
        }                                                                                          \
    } while (0)
    constexpr int DMA_PER_STAGE = 2 * C::DMA_PER_SIDE;  // per thread (per wave instruction count)

    f32x16_t acc[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int e = 0; e < 16; e++) acc[rb][cb][e] = 0.0f;

    // Fragment registers, two sets.  bf16: hi and lo 8-element vectors per
    // block; fp32: two chunks (8 consecutive k) per block for lane half h.
    u32x4 fa0[4][2], fb0[4][2], fa1[4][2], fb1[4][2];
    // logical chunks a lane reads: bf16 (hi h, lo h) = (h, 2 + h); bf16 plain (h);
    // fp32 (2h, 2h + 1)
    const int c0 = (PREC == 0) ? 2 * half : half;
    const int c1 = (PREC == 0) ? 2 * half + 1 : 2 + half;
#define BNA_LDFRAG(FA, FB, STAGE_IDX)                                                              \
    do {                                                                                           \
        const unsigned char *slot_ = smem + ((STAGE_IDX) % NSLOT) * C::SLOT;                       \
        _Pragma("unroll") for (int blk = 0; blk < 4; blk++) {                                      \
            const int ra = wr0 + blk * 32 + lrow, rbw = wc0 + blk * 32 + lrow;                     \
            FA[blk][0] = *(const u32x4 *)(slot_ + ra * C::UNIT + ((c0 ^ swz<C::CH>(ra)) << 4));    \
            FB[blk][0] = *(const u32x4 *)(slot_ + C::PANEL + rbw * C::UNIT + ((c0 ^ swz<C::CH>(rbw)) << 4)); \
            if constexpr (PREC != 1) {                                                             \
                FA[blk][1] = *(const u32x4 *)(slot_ + ra * C::UNIT + ((c1 ^ swz<C::CH>(ra)) << 4)); \
                FB[blk][1] = *(const u32x4 *)(slot_ + C::PANEL + rbw * C::UNIT + ((c1 ^ swz<C::CH>(rbw)) << 4)); \
            }                                                                                      \
        }                                                                                          \
    } while (0)
#define BNA_MFMA(FA, FB)                                                                           \
    do {                                                                                           \
        if constexpr (PREC == 0) {                                                                 \
            _Pragma("unroll") for (int q = 0; q < 8; q++)                                          \
                _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                   \
                    _Pragma("unroll") for (int cb = 0; cb < 4; cb++) {                             \
                        const float av = __builtin_bit_cast(f32x4_t, FA[rb][q >> 2])[q & 3];       \
                        const float bv = __builtin_bit_cast(f32x4_t, FB[cb][q >> 2])[q & 3];       \
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[rb][cb], 0, 0, 0); \
                    }                                                                              \
        } else {                                                                                   \
            _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                       \
                _Pragma("unroll") for (int cb = 0; cb < 4; cb++)                                   \
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                         \
                        __builtin_bit_cast(bf16x8_t, FA[rb][0]), __builtin_bit_cast(bf16x8_t, FB[cb][0]), \
                        acc[rb][cb], 0, 0, 0);                                                     \
            if constexpr (PREC == 2) {                                                             \
                _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                   \
                    _Pragma("unroll") for (int cb = 0; cb < 4; cb++)                               \
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                     \
                            __builtin_bit_cast(bf16x8_t, FA[rb][0]), __builtin_bit_cast(bf16x8_t, FB[cb][1]), \
                            acc[rb][cb], 0, 0, 0);                                                 \
                _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                   \
                    _Pragma("unroll") for (int cb = 0; cb < 4; cb++)                               \
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                     \
                            __builtin_bit_cast(bf16x8_t, FA[rb][1]), __builtin_bit_cast(bf16x8_t, FB[cb][0]), \
                            acc[rb][cb], 0, 0, 0);                                                 \
            }                                                                                      \
        }                                                                                          \
    } while (0)

    // Prologue: stages 0..3 in flight; wait for stage 0, read its fragments.
    BNA_DMA(0);
    BNA_DMA(1);
    BNA_DMA(2);
    BNA_DMA(3);
    BNA_WAIT_VMCNT(3 * DMA_PER_STAGE);
    __builtin_amdgcn_s_barrier();
    BNA_LDFRAG(fa0, fb0, 0);

    // Iteration kt: fragments of stage kt are in set (kt & 1).  Wait for stage
    // kt + 1 (two later stages may stay in flight), barrier (every wave has its
    // stage-kt fragments, so slot kt % 4 is free), refill that slot with stage
    // kt + 4, read stage kt + 1's fragments, then multiply stage kt.  L_pad is a
    // multiple of 32, so nk is even and the loop runs whole pairs of stages.
    for (int kt = 0; kt < nk; kt += 2) {
        BNA_WAIT_VMCNT(2 * DMA_PER_STAGE);
        __builtin_amdgcn_s_barrier();
        BNA_DMA(kt + 4);
        BNA_LDFRAG(fa1, fb1, kt + 1);
        BNA_MFMA(fa0, fb0);
        BNA_WAIT_VMCNT(2 * DMA_PER_STAGE);
        __builtin_amdgcn_s_barrier();
        BNA_DMA(kt + 5);
        BNA_LDFRAG(fa0, fb0, kt + 2);
        BNA_MFMA(fa1, fb1);
    }
    // drain the ring before the epilogue reuses the LDS
    BNA_WAIT_VMCNT(0);
    __syncthreads();
#undef BNA_DMA
#undef BNA_LDFRAG
#undef BNA_MFMA

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
        int *rowoff = (int *)(colbits + TILE * 8);   // [256] first slot of each tile row
        int *coloff = rowoff + TILE;                 // [256] first slot of each tile column
        int *wsum = coloff + TILE;                   // [4]
        int *wsum2 = wsum + 4;                       // [4]
        unsigned long long *sbase = (unsigned long long *)(wsum2 + 4);
        KeptPair *lists = (KeptPair *)(smem + EPI_LIST_OFF);  // [4 waves][LIST_LDS]
        for (int k = tid; k < TILE * 8 * 2; k += CT_THREADS) rowbits[k] = 0;
        __syncthreads();

        WaveList wl{(lds_u64 *)(lists + wave * LIST_LDS), nullptr, 0, false};
        if (ec.diag || ec.gi0 + TILE > sd.N || ec.gj0 + TILE > sd.N)
            BNA_FOR_BLOCKS(epi_mark, ec, rowbits, colbits, wl, sd.ovf, sd.ovf_count, sd.ovf_cap);
        else
            BNA_FOR_BLOCKS(epi_mark_inner, ec, rowbits, colbits, wl, sd.ovf, sd.ovf_count, sd.ovf_cap);
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

        // fragment tables, [k][row] so that a tile's writes are contiguous
        const int N = sd.N;
        if (ec.gi0 + tid < N) {
            const int64_t f = (int64_t)(J + 1) * N + ec.gi0 + tid;
            sd.frag_orig[f] = rc;
            sd.frag_live[f] = rc;
            sd.frag_off[f] = base + ro;
        }
        if (ec.gj0 + tid < N) {
            const int64_t f = (int64_t)I * N + ec.gj0 + tid;
            sd.frag_orig[f] = cc;
            sd.frag_live[f] = cc;
            sd.frag_off[f] = base + total + co;
        }
        if (wl.n == 0 || wl.lost) return;
        if (wl.n > LIST_LDS) __builtin_amdgcn_s_waitcnt(0);  // overflow entries stored by this wave
        const float sure = sd.sure_thr;
        for (int k = lane; k < wl.n; k += 64) {
            uint64_t raw;
            if (k < LIST_LDS)
                raw = wl.lds[k];
            else
                raw = __builtin_nontemporal_load(wl.ovf + (k - LIST_LDS));
            const uint32_t lo32 = (uint32_t)raw;
            const int lr = lo32 & 0xffff, lc = lo32 >> 16;
            const float v = __builtin_bit_cast(float, (uint32_t)(raw >> 32));
            int p = rowoff[lr], q = coloff[lc];
            const int rwd = lc >> 5, cwd = lr >> 5;
            for (int w = 0; w < rwd; w++) p += __popc(rowbits[lr * 8 + w]);
            for (int w = 0; w < cwd; w++) q += __popc(colbits[lc * 8 + w]);
            p += __popc(rowbits[lr * 8 + rwd] & ((1u << (lc & 31)) - 1u));
            q += __popc(colbits[lc * 8 + cwd] & ((1u << (lr & 31)) - 1u));
            const int64_t ps = base + p, qs = base + total + q;
            if (fits) {
                sd.entries[ps] = Entry{ec.gj0 + lc, v};
                sd.entries[qs] = Entry{ec.gi0 + lr, v};
            }
            if (!(v > sure)) {
                const unsigned long long ci = atomicAdd(sd.cand_count, 1ull);
                if ((int64_t)ci < sd.cand_cap) sd.cands[ci] = Candidate{ec.gi0 + lr, ec.gj0 + lc, ps, qs, v, 0};
            }
        }
    }
}

template <int PREC, int HIST>
static hipError_t launch_one(const SubjDesc *subj_dev, const TileItem *items, int64_t n_items, hipStream_t st)
{
    using C = CorrCfg<PREC>;
    constexpr size_t epi = (size_t)EPI_LIST_OFF + 4 * LIST_LDS * sizeof(KeptPair);
    constexpr size_t hst = (size_t)HIST_BINS * 4;
    constexpr size_t ring = (size_t)NSLOT * C::SLOT;
    constexpr size_t shm = ring > epi ? (ring > hst ? ring : hst) : (epi > hst ? epi : hst);
    static std::atomic<int> attr_dev{-1};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (attr_dev.load() != dev) {
        e = hipFuncSetAttribute((const void *)corr_tile_kernel<PREC, HIST>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        attr_dev.store(dev);
    }
    const int64_t grid = (n_items + 7) / 8 * 8;
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
