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
// Persistent launch: one workgroup per CU claims tiles from the tile table until
// it is drained.  The table is split into eight contiguous ranges, one per XCD
// (workgroup b runs on XCD b % 8), so tiles that share operand panels run on the
// same L2; a workgroup whose range is drained takes tiles from the others.  Each
// workgroup knows its next tile one tile ahead, and the operand stream runs on
// across the tile boundary: the last four stages of a tile fetch the next
// tile's first four stages, so those loads land while the epilogue runs.  A batch of
// subjects (different N, L, thresholds) shares one launch: every tile item
// carries its subject's operand panel and stage count.
//
// Work split: a 256 x 256 tile per workgroup of 4 waves; each wave owns a
// 128 x 128 sub-tile held as 4 x 4 MFMA 32x32 accumulator blocks (256 AGPRs).
//
// Operand panels are k-blocked: for each block of 16 time points kb and each
// voxel row r there is one UNIT-byte record at ((kb * N_pad) + r) * UNIT, so
// one stage of a 256-row panel is a contiguous 256 * UNIT bytes.  Stages move
// global -> LDS with global_load_lds_dwordx4 (no registers in between) into a
// ring of 4 slots.  Stage p's fragments are read from LDS into registers while
// stage p - 1 is multiplied, and the load for stream position p + 4 is issued
// when stage p's multiply starts, so every load has about three stages of MFMA
// work to arrive in and no LDS read waits in front of an MFMA.  (The dense
// accumulation pass keeps the one-set form: position p + 3 while p is read and
// multiplied.)
//
// Precisions (template PREC), all accumulating in FP32:
//   0 FP32    v_mfma_f32_32x32x2_f32,   UNIT 64: z[16] fp32
//   1 BF16    v_mfma_f32_32x32x16_bf16, UNIT 32: hi[16]
//   2 BF16X3  v_mfma_f32_32x32x16_bf16, UNIT 64: hi[16] lo[16]; hi*hi + hi*lo + lo*hi
//
// Epilogue (in an LDS region beside the ring): each kept pair (r > keep_thr)
// sets a bit in a tile-row and a tile-column bitmap; per-row and per-column
// popcounts, two block scans and one global atomic give every pair its slot in
// both lists.  Kept pairs are recorded in a per-wave LDS list that continues, for
// dense tiles, in a per-wave spill area in global memory.  Pairs with
// r <= sure_thr are also queued as candidates for the FP64 recheck.  In
// histogram mode the epilogue adds every valid pair's bin to an LDS histogram
// that lives across tiles and is flushed when the subject changes.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>

namespace bna {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) void gvoid;
// Device-memory pointers of the subject descriptor, used as global-address-space
// pointers: a generic (flat) access may touch LDS, so the compiler would make it
// wait for the stage loads in flight to the LDS ring first.
typedef __attribute__((address_space(1))) int32_t g_i32;
typedef __attribute__((address_space(1))) int64_t g_i64;
typedef __attribute__((address_space(1))) unsigned long long g_u64;
typedef __attribute__((address_space(1))) float g_f32;
static_assert(sizeof(Entry) == 8 && sizeof(Candidate) == 32, "entry and candidate records are stored as 64-bit words");
__device__ __forceinline__ uint64_t pack2(uint32_t lo, uint32_t hi) { return (uint64_t)lo | ((uint64_t)hi << 32); }
__device__ __forceinline__ unsigned long long gatomic_add(unsigned long long *p, unsigned long long v)
{
    return __hip_atomic_fetch_add((g_u64 *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
typedef __attribute__((address_space(3))) void lvoid;

constexpr int CT_THREADS = 256;
constexpr int NSLOT = 4;                   // LDS ring of k16 stages (power of two)
constexpr int LDS_MAX = 160 * 1024;        // gfx950 LDS per workgroup
constexpr int CTL_BYTES = 64;              // tile-claim broadcast words at the top of LDS
constexpr int EPI_FIXED = 16384 + 2048 + 64 + 4096;  // bitmaps, offsets, scan partials, word prefixes

template <int PREC>
struct CorrCfg {
    static constexpr int UNIT = (PREC == 1) ? 32 : 64;   // bytes per (k16 block, row)
    static constexpr int CH = UNIT / 16;                 // 16-byte chunks per row
    static constexpr int PANEL = TILE * UNIT;            // bytes of one side of a stage
    static constexpr int SLOT = 2 * PANEL;               // A + B
    static constexpr int DMA_PER_SIDE = PANEL / 1024 / 4;  // 1 KB wave instructions per wave per side
    static constexpr int DPS = 2 * DMA_PER_SIDE;         // DMA instructions per wave per stage
    static constexpr int RING = NSLOT * SLOT;            // epilogue region starts here
    static constexpr int LIST_RAW = (LDS_MAX - CTL_BYTES - RING - EPI_FIXED) / (4 * 8);
    static constexpr int LIST_LDS = (LIST_RAW > 2048 ? 2048 : LIST_RAW) & ~7;  // kept pairs per wave
};
static_assert(CorrCfg<2>::LIST_LDS >= 256, "epilogue lists do not fit beside the ring");
static_assert(LPAD_MIN >= 4 * KPAD, "the operand stream fetches four stages of the next tile");
static_assert(CorrCfg<0>::RING + EPI_FIXED + CTL_BYTES <= LDS_MAX, "LDS layout");
static_assert(HIST_BINS * 4 <= EPI_FIXED + 4 * 8 * CorrCfg<2>::LIST_LDS, "histogram region");

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
// lets the slot be refilled.  The stage loop pairs it with a bare s_barrier: a
// fence there would make the compiler wait for every stage load in flight.
#define BNA_WAIT_VMCNT(n) \
    __builtin_amdgcn_s_waitcnt((((n) & 0xf) | (((n) >> 4) << 14)) | (0x7 << 4))

// Workgroup barrier for the epilogue's LDS traffic: the wave's LDS operations
// complete (lgkmcnt 0) and the compiler moves no memory access across it.
// Outstanding global loads and stores, the next tile's operand stream among
// them, are not waited for (an LDS fence would wait for that stream, whose LDS
// writes are tracked by vmcnt; the stream's slots are disjoint from the
// epilogue's LDS and are waited for before the next tile reads them).
__device__ __forceinline__ void lds_sync() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), vmcnt and expcnt unconstrained
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__device__ __forceinline__ int64_t uni64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Next tile for a workgroup homed on XCD `xcd`: its own range first, then the
// others in turn.  Every index of every range is handed out exactly once.
__device__ int64_t claim_tile(unsigned int *claim, int64_t n_items, int64_t per, int xcd)
{
    for (int r = 0; r < 8; r++) {
        const int x = (xcd + r) & 7;
        const int64_t lo = (int64_t)x * per;
        const int64_t cnt = n_items - lo < per ? n_items - lo : per;
        if (cnt <= 0) continue;
        if ((int64_t)__hip_atomic_load(&claim[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= cnt) continue;
        const unsigned int t = atomicAdd(&claim[x], 1u);
        if ((int64_t)t < cnt) return lo + (int64_t)t;
    }
    return -1;
}

// One side of a stage: NQ wave instructions of 1 KB each from one global
// address and one LDS address; the immediate offset q * 1024 advances both.
template <int NQ>
__device__ __forceinline__ void dma_side(const unsigned char *g, unsigned char *l)
{
    static_assert(NQ == 2 || NQ == 4, "stage side of 2 or 4 KB per wave");
    __builtin_amdgcn_global_load_lds((gvoid *)g, (lvoid *)l, 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gvoid *)g, (lvoid *)l, 16, 1024, 0);
    if constexpr (NQ == 4) {
        __builtin_amdgcn_global_load_lds((gvoid *)g, (lvoid *)l, 16, 2048, 0);
        __builtin_amdgcn_global_load_lds((gvoid *)g, (lvoid *)l, 16, 3072, 0);
    }
}

// Block-wide exclusive scan of one int per thread (256 threads).
__device__ __forceinline__ int block_excl_scan(int v, int *wsum, int &total, int lane, int w) {
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    lds_sync();
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
    int N, gi0, gj0, wr0, wc0, half, lrow, lane;
    bool diag;
    float keep;
};

// Records are stored packed in one 64-bit word (tile row | tile column << 16
// in the low half, r bits in the high half).  The first `cap` records of a wave
// go to LDS, the rest (dense tiles only) to the wave's own spill area in global
// memory, which holds the remainder of a whole 128 x 128 sub-tile.
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef __attribute__((address_space(1))) uint64_t glb_u64;
constexpr int WAVE_PAIRS = 128 * 128;
struct WaveList {
    lds_u64 *lds;
    glb_u64 *spill;
    int cap;  // LDS entries
    int n;    // kept pairs of the wave so far (wave-uniform)
};

__device__ __forceinline__ bool epi_in(const EpiCtx &c, int lr, int lc) {
    return (c.gi0 + lr < c.N) && (c.gj0 + lc < c.N) && (!c.diag || lr < lc);
}

// Fisher z of one r as the group average takes it (CorMat.cpp:36-42): r == 1 is
// moved to the float below 1 - 1e-6, the quotient and its log are float, the
// half is taken in double and the result stored as float.
__device__ __forceinline__ float fisher_z(float r) {
    if (r == 1.0f) r = (float)((double)r - 1e-6);
    return (float)(0.5 * (double)logf((1.0f + r) / (1.0f - r)));
}

// Accumulation pass: every pair of the block whose r is a number (a NaN from an
// invalid voxel is not) adds r, or its Fisher z, to its slot of the packed upper
// triangle, row-major with i < j (CorMat.cpp:851-866).  The reference adds r only
// when it lies in [-1, 1]; the r of two valid voxels always does, so an MFMA
// value just outside it (near-identical series) is clamped to the range instead
// of being dropped, which keeps every term within the precision's error band of
// the FP64 value.  For Fisher sums the pairs with |r| > flag_r in some subject
// are marked in a bitmap (their z is too steep for the group's error bound).
// Lanes 0-31 and 32-63 touch two runs of 32 neighbouring slots.  Each pair
// belongs to one tile, so the read-modify-write needs no atomic.
// A tile inside the matrix and off the diagonal updates every slot of the
// block without a branch (a NaN adds +0, which leaves the sum as it is), so
// its sixteen loads can be in flight together; edge and diagonal tiles test
// each element first.
__device__ __forceinline__ float accum_term(float v, int fisher) {
    if (!(v == v)) return 0.0f;
    const float r = fminf(fmaxf(v, -1.0f), 1.0f);
    return fisher ? fisher_z(r) : r;
}
__device__ __forceinline__ void accum_flag(const EpiCtx &c, uint32_t *flags, float flag_r, float v, int64_t p) {
    if (flags && fabsf(v) > flag_r)
        __hip_atomic_fetch_or((__attribute__((address_space(1))) uint32_t *)flags + (p >> 5), 1u << (p & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int RB, int CB>
__device__ __forceinline__ void epi_accum(const f32x16_t &a, const EpiCtx &c, float *dense, int fisher,
                                          uint32_t *flags, float flag_r) {
    const int lc = c.wc0 + CB * 32 + c.lrow;
    const int64_t j = c.gj0 + lc;
    const bool inner = !c.diag && c.gi0 + TILE <= c.N && c.gj0 + TILE <= c.N;
    if (inner) {
        float add[16];
        int64_t pos[16];
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int64_t i = c.gi0 + c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
            add[e] = accum_term(a[e], fisher);
            pos[e] = i * c.N - (i + 1) * i / 2 + j - i - 1;
        }
        float old[16];
#pragma unroll
        for (int e = 0; e < 16; e++) old[e] = ((g_f32 *)dense)[pos[e]];
#pragma unroll
        for (int e = 0; e < 16; e++) ((g_f32 *)dense)[pos[e]] = old[e] + add[e];
        if (flags) {
#pragma unroll
            for (int e = 0; e < 16; e++) accum_flag(c, flags, flag_r, a[e], pos[e]);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        const float v = a[e];
        if (epi_in(c, lr, lc) && v == v) {
            const int64_t i = c.gi0 + lr;
            const int64_t p = i * c.N - (i + 1) * i / 2 + j - i - 1;
            ((g_f32 *)dense)[p] += accum_term(v, fisher);
            accum_flag(c, flags, flag_r, v, p);
        }
    }
}

__device__ __forceinline__ void epi_hist(const f32x16_t &a, int RB, int CB, const EpiCtx &c, uint32_t *hist) {
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

__device__ __forceinline__ void epi_record(WaveList &wl, uint64_t m, uint64_t lt, int lr, int lc, float v,
                                           uint32_t *rowbits, uint32_t *colbits)
{
    atomicOr(&rowbits[lr * 8 + (lc >> 5)], 1u << (lc & 31));
    atomicOr(&colbits[lc * 8 + (lr >> 5)], 1u << (lr & 31));
    const int k = wl.n + __popcll(m & lt);
    const uint64_t rec = (uint64_t)((uint32_t)lr | ((uint32_t)lc << 16)) |
                         ((uint64_t)__builtin_bit_cast(uint32_t, v) << 32);
    if (k < wl.cap)
        wl.lds[k] = rec;
    else
        wl.spill[k - wl.cap] = rec;
}

// The single pass over one accumulator block of a tile that lies inside the
// matrix and off the diagonal, where every element is a pair to test: marks
// the kept pairs in both bitmaps and appends them to the wave's list at
// ballot-compacted positions.  The block is tested four elements (one row
// quad) at a time: a quad none of whose elements exceeds the threshold costs
// three maxima, one compare and one ballot (a NaN never wins the maximum and is
// never kept); inside a quad only the elements some lane keeps take the
// recording branch.
template <int RB, int CB>
__device__ __forceinline__ void epi_mark_inner(const f32x16_t &a, const EpiCtx &c, uint32_t *rowbits,
                                               uint32_t *colbits, WaveList &wl)
{
    const int lc = c.wc0 + CB * 32 + c.lrow;
    const uint64_t lt = (1ull << c.lane) - 1ull;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float mx = fmaxf(fmaxf(a[4 * g], a[4 * g + 1]), fmaxf(a[4 * g + 2], a[4 * g + 3]));
        if (__builtin_expect(__ballot(mx > c.keep) == 0, 1)) continue;  // kept pairs are rare: the
                                                                         // recording code is laid out cold
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int e = 4 * g + q;
            const bool kp = a[e] > c.keep;
            const uint64_t m = __ballot(kp);
            if (m == 0) continue;
            if (kp) epi_record(wl, m, lt, c.wr0 + RB * 32 + g * 8 + c.half * 4 + q, lc, a[e], rowbits, colbits);
            wl.n += __popcll(m);
        }
    }
}

// The same for a tile on the diagonal or at the matrix edge, where each element
// is first checked to be a pair (row < column on the diagonal, both inside N).
template <int RB, int CB>
__device__ __forceinline__ void epi_mark_edge(const f32x16_t &a, const EpiCtx &c, uint32_t *rowbits,
                                              uint32_t *colbits, WaveList &wl)
{
    const int lc = c.wc0 + CB * 32 + c.lrow;
    const uint64_t lt = (1ull << c.lane) - 1ull;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int lr = c.wr0 + RB * 32 + (e >> 2) * 8 + c.half * 4 + (e & 3);
        const bool kp = a[e] > c.keep && epi_in(c, lr, lc);
        const uint64_t m = __ballot(kp);
        if (__builtin_expect(m != 0, 0)) {
            if (kp) epi_record(wl, m, lt, lr, lc, a[e], rowbits, colbits);
            wl.n += __popcll(m);
        }
    }
}

// Slots of pair (lr, lc) in the tile's row-view and column-view lists.  Byte w
// of rowpre[r] (colpre[c]) is the count of kept pairs in words 0 .. w - 1 of
// row r's (column c's) bitmap, so a slot is three LDS reads and a popcount.
struct SlotCtx {
    const uint32_t *rowbits, *colbits;
    const int *rowoff, *coloff;
    const uint64_t *rowpre, *colpre;
    int64_t base, total;
};
__device__ __forceinline__ void pair_slots(const SlotCtx &s, int lr, int lc, int64_t &ps, int64_t &qs, int &qw)
{
    const int rwd = lc >> 5, cwd = lr >> 5;
    const int pw = (int)((s.rowpre[lr] >> (8 * rwd)) & 0xff) +
                   __popc(s.rowbits[lr * 8 + rwd] & ((1u << (lc & 31)) - 1u));
    const int qp = (int)((s.colpre[lc] >> (8 * cwd)) & 0xff) +
                   __popc(s.colbits[lc * 8 + cwd] & ((1u << (lr & 31)) - 1u));
    ps = s.base + s.rowoff[lr] + pw;
    qs = s.base + s.total + s.coloff[lc] + qp;
    qw = qp;  // position within column lc's list
}

// Writes one kept pair's two entries and, for band pairs, its candidate record
// (one counter update per wave and call, slots by ballot rank).
__device__ __forceinline__ void emit_pair(const SubjDesc &sd, const EpiCtx &c, const SlotCtx &s, bool fits, bool act,
                                          int lr, int lc, float v)
{
    int64_t ps = 0, qs = 0;
    if (act) {
        int qw = 0;
        pair_slots(s, lr, lc, ps, qs, qw);
        if (fits) {
            // where the value refinement finds the pair's column-view entry
            if (sd.mirror) ((__attribute__((address_space(1))) uint8_t *)sd.mirror)[ps] = (uint8_t)qw;
            // Entry {int32 col; float r} as one 64-bit word
            ((g_u64 *)sd.entries)[ps] = pack2((uint32_t)(c.gj0 + lc), __builtin_bit_cast(uint32_t, v));
            ((g_u64 *)sd.entries)[qs] = pack2((uint32_t)(c.gi0 + lr), __builtin_bit_cast(uint32_t, v));
        }
    }
    const bool cand = act && !(v > sd.sure_thr);
    const uint64_t cm = __ballot(cand);
    if (cm) {
        const int lane = c.lane;
        const int leader = __builtin_ctzll(cm);
        unsigned long long cb = 0;
        if (lane == leader) cb = gatomic_add(sd.cand_count, (unsigned long long)__popcll(cm));
        cb = __shfl(cb, leader, 64);
        if (cand) {
            const unsigned long long ci = cb + (unsigned long long)__popcll(cm & ((1ull << lane) - 1ull));
            if ((int64_t)ci < sd.cand_cap) {  // Candidate {i, j, slot_row, slot_col, r, pad} as four words
                g_u64 *q = (g_u64 *)sd.cands + 4 * (int64_t)ci;
                q[0] = pack2((uint32_t)(c.gi0 + lr), (uint32_t)(c.gj0 + lc));
                q[1] = (uint64_t)ps;
                q[2] = (uint64_t)qs;
                q[3] = pack2(__builtin_bit_cast(uint32_t, v), 0u);
            }
        }
    }
}

// One accumulator block at a time: the block is copied to VGPRs behind an
// opaque barrier, so the compiler cannot interleave (and keep live) the copies of
// several blocks; without it the epilogue spills.
#define BNA_BLOCK_T(F, RB, CB, ...)                                                                \
    do {                                                                                           \
        f32x16_t v_ = acc[RB][CB];                                                                 \
        asm volatile("" : "+v"(v_));                                                               \
        F<RB, CB>(v_, __VA_ARGS__);                                                                \
    } while (0)
#define BNA_FOR_BLOCKS_T(F, ...)                                                                   \
    do {                                                                                           \
        BNA_BLOCK_T(F, 0, 0, __VA_ARGS__); BNA_BLOCK_T(F, 0, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 0, 2, __VA_ARGS__); BNA_BLOCK_T(F, 0, 3, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 1, 0, __VA_ARGS__); BNA_BLOCK_T(F, 1, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 1, 2, __VA_ARGS__); BNA_BLOCK_T(F, 1, 3, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 2, 0, __VA_ARGS__); BNA_BLOCK_T(F, 2, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 2, 2, __VA_ARGS__); BNA_BLOCK_T(F, 2, 3, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 3, 0, __VA_ARGS__); BNA_BLOCK_T(F, 3, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 3, 2, __VA_ARGS__); BNA_BLOCK_T(F, 3, 3, __VA_ARGS__);                      \
    } while (0)
#define BNA_FOR_BLOCKS(F, ...)                                                                     \
    do {                                                                                           \
        F(acc[0][0], 0, 0, __VA_ARGS__); F(acc[0][1], 0, 1, __VA_ARGS__);                          \
        F(acc[0][2], 0, 2, __VA_ARGS__); F(acc[0][3], 0, 3, __VA_ARGS__);                          \
        F(acc[1][0], 1, 0, __VA_ARGS__); F(acc[1][1], 1, 1, __VA_ARGS__);                          \
        F(acc[1][2], 1, 2, __VA_ARGS__); F(acc[1][3], 1, 3, __VA_ARGS__);                          \
        F(acc[2][0], 2, 0, __VA_ARGS__); F(acc[2][1], 2, 1, __VA_ARGS__);                          \
        F(acc[2][2], 2, 2, __VA_ARGS__); F(acc[2][3], 2, 3, __VA_ARGS__);                          \
        F(acc[3][0], 3, 0, __VA_ARGS__); F(acc[3][1], 3, 1, __VA_ARGS__);                          \
        F(acc[3][2], 3, 2, __VA_ARGS__); F(acc[3][3], 3, 3, __VA_ARGS__);                          \
    } while (0)

// Adds a workgroup's LDS histogram to the subject's global one and clears it.
__device__ __forceinline__ void hist_flush(uint32_t *hist, unsigned long long *dst, int tid)
{
    for (int k = tid; k < HIST_BINS; k += CT_THREADS) {
        const uint32_t h = hist[k];
        if (h) {
            gatomic_add(dst + k, (unsigned long long)h);
            hist[k] = 0;
        }
    }
}

template <int PREC, int HIST>
__global__ __launch_bounds__(CT_THREADS, 1) void corr_tile_kernel(const SubjDesc *__restrict__ subjs,
                                                                  const TileItem *__restrict__ items,
                                                                  int64_t n_items, unsigned int *__restrict__ claim,
                                                                  uint64_t *__restrict__ spill)
{
    using C = CorrCfg<PREC>;
    extern __shared__ __align__(1024) unsigned char smem[];
    unsigned char *const epi = smem + C::RING;
    int64_t *const ctl = (int64_t *)(smem + LDS_MAX - CTL_BYTES);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int swave = __builtin_amdgcn_readfirstlane(wave);  // the same value, held in a scalar register
    const int wr0 = (wave >> 1) * 128;
    const int wc0 = (wave & 1) * 128;
    const int lrow = lane & 31;
    const int half = lane >> 5;
    const int xcd = blockIdx.x & 7;
    const int64_t per_xcd = (n_items + 7) / 8;

    if (tid == 0) {
        const int64_t t0 = claim_tile(claim, n_items, per_xcd, xcd);
        ctl[0] = t0;
        ctl[1] = t0 >= 0 ? claim_tile(claim, n_items, per_xcd, xcd) : -1;
    }
    if constexpr (HIST == 1) {
        for (int k = tid; k < HIST_BINS; k += CT_THREADS) ((uint32_t *)epi)[k] = 0;
    }
    __syncthreads();
    int64_t cur = uni64(ctl[0]);
    int64_t nxt = uni64(ctl[1]);
    if (cur < 0) return;
    TileItem ic = items[cur];
    TileItem in = nxt >= 0 ? items[nxt] : ic;

    // DMA addressing.  A wave instruction fills one contiguous KB of LDS, lane l
    // the 16 bytes at l * 16: row l / CH of the KB's rows, physical chunk l % CH,
    // which holds logical chunk (l % CH) ^ swz(row).  Wave w fills rows
    // [w * 64, w * 64 + 64) of each side.
    // Instruction q covers rows w * 64 + q * (1024 / UNIT) + lane / CH; those row
    // offsets are multiples of 16, so the swizzle does not depend on q and the
    // byte offset is dsrc + q * 1024.  The LDS destination advances by the same
    // 1024 bytes, so the four instructions of a side share one global address and
    // one M0 value and differ only in the immediate offset, which the instruction
    // applies to both addresses.
    const uint32_t dsrc = (uint32_t)((wave * (TILE / 4) + lane / C::CH) * C::UNIT +
                                     (((lane % C::CH) ^ swz<C::CH>(lane / C::CH)) << 4));
    int g = 0;  // stream position of the current tile's stage 0; stage s lives in slot (g + s) % NSLOT

    // Stream position s relative to the current tile: its own stage s, then the
    // next tile's stages (s - nk < 4 <= the next tile's nk) from xA/xB with
    // stride xK, or, after the last tile, a repeat of the last stage (stride 0)
    // so every wave issues the same count.
#define BNA_DMA(S)                                                                                 \
    do {                                                                                           \
        const int s_ = (S);                                                                        \
        const bool own_ = s_ < nk;                                                                 \
        const int64_t o_ = own_ ? s_ * ic.kstride : (s_ - nk) * xK;                                \
        const unsigned char *a_ = (own_ ? cA : xA) + o_;                                           \
        const unsigned char *b_ = (own_ ? cB : xB) + o_;                                           \
        unsigned char *dA_ = smem + ((g + s_) & (NSLOT - 1)) * C::SLOT + swave * C::DMA_PER_SIDE * 1024; \
        const unsigned char *ga_ = a_ + dsrc;                                                      \
        const unsigned char *gb_ = b_ + dsrc;                                                      \
        dma_side<C::DMA_PER_SIDE>(ga_, dA_);                                                       \
        dma_side<C::DMA_PER_SIDE>(gb_, dA_ + C::PANEL);                                            \
    } while (0)

    // Fragment registers: bf16 hi and lo 8-element vectors per block; fp32 two
    // chunks (8 consecutive k) per block for lane half h.  Two sets: stage kt + 1
    // is read into one while stage kt's MFMAs run from the other, so the LDS
    // reads never stall the matrix pipe.  (The accumulation pass keeps one set:
    // with a second one its epilogue spills inside the stage loop.)
    u32x4 fa0[4][2], fb0[4][2], fa1[4][2], fb1[4][2];
    constexpr int AHEAD = (HIST == 2) ? 3 : 4;  // stream positions in flight ahead of the stage being read
    // logical chunks a lane reads: bf16 (hi h, lo h) = (h, 2 + h); bf16 plain (h);
    // fp32 (2h, 2h + 1)
    const int c0 = (PREC == 0) ? 2 * half : half;
    const int c1 = (PREC == 0) ? 2 * half + 1 : 2 + half;
    // Row wr0 + 32 blk + lrow of a slot; block offsets are multiples of 32 rows,
    // so the swizzle is that of lrow and block blk sits 32 * blk * UNIT further.
    const int sw = swz<C::CH>(lrow);
    const uint32_t fA0 = (uint32_t)((wr0 + lrow) * C::UNIT + ((c0 ^ sw) << 4));
    const uint32_t fA1 = (uint32_t)((wr0 + lrow) * C::UNIT + ((c1 ^ sw) << 4));
    const uint32_t fB0 = (uint32_t)(C::PANEL + (wc0 + lrow) * C::UNIT + ((c0 ^ sw) << 4));
    const uint32_t fB1 = (uint32_t)(C::PANEL + (wc0 + lrow) * C::UNIT + ((c1 ^ sw) << 4));
#define BNA_LDFRAG(FA, FB, STAGE_IDX)                                                              \
    do {                                                                                           \
        const unsigned char *slot_ = smem + ((g + (STAGE_IDX)) & (NSLOT - 1)) * C::SLOT;           \
        _Pragma("unroll") for (int blk = 0; blk < 4; blk++) {                                      \
            FA[blk][0] = *(const u32x4 *)(slot_ + fA0 + blk * 32 * C::UNIT);                       \
            FB[blk][0] = *(const u32x4 *)(slot_ + fB0 + blk * 32 * C::UNIT);                       \
            if constexpr (PREC != 1) {                                                             \
                FA[blk][1] = *(const u32x4 *)(slot_ + fA1 + blk * 32 * C::UNIT);                   \
                FB[blk][1] = *(const u32x4 *)(slot_ + fB1 + blk * 32 * C::UNIT);                   \
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

    // Issue order of one stage after its barrier: an MFMA first (the fragments
    // are already in registers), then each stage load and each LDS read placed
    // behind an MFMA of its own, so their issue cycles fall in the shadow of
    // MFMAs already running instead of leaving the matrix pipe idle.
    constexpr int NMF = (PREC == 0) ? 128 : (PREC == 1 ? 16 : 48);  // MFMAs per stage
    constexpr int NRD = (PREC == 1) ? 8 : 16;                        // LDS fragment reads per stage
#define BNA_INTERLEAVE()                                                                           \
    do {                                                                                           \
        _Pragma("unroll") for (int i_ = 0; i_ < C::DPS; i_++) {                                    \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                     \
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);                                     \
        }                                                                                          \
        _Pragma("unroll") for (int i_ = 0; i_ < NRD; i_++) {                                       \
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                                     \
            __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);                                     \
        }                                                                                          \
        __builtin_amdgcn_sched_group_barrier(0x008, NMF - C::DPS - NRD, 0);                        \
    } while (0)

    // Prologue of the first tile: stream positions 0 .. AHEAD - 1.
    {
        const int nk = ic.nk;
        const unsigned char *cA = (const unsigned char *)ic.panel + (int64_t)ic.I * C::PANEL;
        const unsigned char *cB = (const unsigned char *)ic.panel + (int64_t)ic.J * C::PANEL;
        const unsigned char *xA = cA, *xB = cB;
        const int64_t xK = 0;
        BNA_DMA(0);
        BNA_DMA(1);
        BNA_DMA(2);
        if constexpr (AHEAD == 4) BNA_DMA(3);
    }

    for (;;) {
        const int nk = ic.nk;
        const unsigned char *cA = (const unsigned char *)ic.panel + (int64_t)ic.I * C::PANEL;
        const unsigned char *cB = (const unsigned char *)ic.panel + (int64_t)ic.J * C::PANEL;
        const unsigned char *xA = nxt >= 0 ? (const unsigned char *)in.panel + (int64_t)in.I * C::PANEL
                                           : cA + (int64_t)(nk - 1) * ic.kstride;
        const unsigned char *xB = nxt >= 0 ? (const unsigned char *)in.panel + (int64_t)in.J * C::PANEL
                                           : cB + (int64_t)(nk - 1) * ic.kstride;
        const int64_t xK = nxt >= 0 ? in.kstride : 0;
        const SubjDesc sd = subjs[ic.subj];  // read now, used by the epilogue

        f32x16_t acc[4][4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[rb][cb][e] = 0.0f;

        // Everything in flight is waited for first: the previous tile's epilogue
        // stores may complete out of order with the stage loads, so only an empty
        // counter proves that the stages issued before them have landed.  (Those
        // stages, this tile's first three, had the epilogue's time to arrive.)
        BNA_WAIT_VMCNT(0);
        // The tile after next is claimed from the home range now; the answer is
        // needed only at the boundary, so the atomic's round trip overlaps the
        // stage loop instead of stalling the epilogue.
        unsigned int pre_t = 0;
        if (tid == 0 && nxt >= 0) pre_t = atomicAdd(&claim[xcd], 1u);
        if constexpr (AHEAD == 4) {
            // Iteration kt: stage kt's fragments are in set kt & 1.  Wait for stage
            // kt + 1 (positions kt + 2 and kt + 3 may stay in flight; this wave's
            // reads of stage kt are complete), barrier (every wave holds stage kt,
            // so its slot is free), refill that slot with position kt + 4, read
            // stage kt + 1 into the other set, multiply stage kt.  The scheduling
            // barrier keeps the MFMAs ahead of the next wait, so the reads have a
            // whole stage of MFMAs to land.
            __builtin_amdgcn_s_barrier();
            BNA_LDFRAG(fa0, fb0, 0);
            int kt = 0;
            for (; kt + 1 < nk; kt += 2) {
                BNA_WAIT_VMCNT(2 * C::DPS);
                __builtin_amdgcn_s_barrier();
                BNA_DMA(kt + 4);
                BNA_LDFRAG(fa1, fb1, kt + 1);
                BNA_MFMA(fa0, fb0);
                BNA_INTERLEAVE();
                __builtin_amdgcn_sched_barrier(0);
                BNA_WAIT_VMCNT(2 * C::DPS);
                __builtin_amdgcn_s_barrier();
                BNA_DMA(kt + 5);
                // read unconditionally (a slot is always inside LDS; past the
                // tile's last stage the values are not used), so the block has no
                // branch and the reads can sit between the MFMAs
                BNA_LDFRAG(fa0, fb0, kt + 2);
                BNA_MFMA(fa1, fb1);
                BNA_INTERLEAVE();
                __builtin_amdgcn_sched_barrier(0);
            }
            if (nk & 1) {  // the last stage, nk - 1, is in set 0
                BNA_WAIT_VMCNT(2 * C::DPS);
                __builtin_amdgcn_s_barrier();
                BNA_DMA(kt + 4);
                BNA_MFMA(fa0, fb0);
            }
        } else {
            // Stage kt: wait for it (positions kt + 1 and kt + 2 may stay in
            // flight), barrier (every wave is done with stage kt - 1, so its slot
            // is free), refill that slot with position kt + 3, read stage kt's
            // fragments and multiply them.
            for (int kt = 0; kt < nk; kt++) {
                BNA_WAIT_VMCNT(2 * C::DPS);
                __builtin_amdgcn_s_barrier();
                BNA_DMA(kt + 3);
                BNA_LDFRAG(fa0, fb0, kt);
                BNA_MFMA(fa0, fb0);
            }
        }

        // ------------------------------------------------------------ epilogue
        // The epilogue's lane coordinates are re-derived from an opaque copy of
        // the thread index so that the compiler does not hoist its ~thousand
        // per-element addresses out of the tile loop (they would not fit in
        // registers across it).
        int etid = tid;
        asm volatile("" : "+v"(etid));
        const int elane = etid & 63, ewave = etid >> 6;
        EpiCtx ec;
        ec.N = sd.N;
        ec.gi0 = ic.I * TILE;
        ec.gj0 = ic.J * TILE;
        ec.wr0 = (ewave >> 1) * 128;
        ec.wc0 = (ewave & 1) * 128;
        ec.half = elane >> 5;
        ec.lrow = elane & 31;
        ec.lane = elane;
        ec.diag = (ic.I == ic.J);
        ec.keep = sd.keep_thr;

        if constexpr (HIST == 1) {
            BNA_FOR_BLOCKS(epi_hist, ec, (uint32_t *)epi);
        } else if constexpr (HIST == 2) {
            BNA_FOR_BLOCKS_T(epi_accum, ec, sd.dense, sd.fisher, sd.flags, sd.flag_r);
        } else {
            uint32_t *rowbits = (uint32_t *)epi;         // [256][8] tile-row bitmaps
            uint32_t *colbits = rowbits + TILE * 8;      // [256][8] tile-column bitmaps
            int *rowoff = (int *)(colbits + TILE * 8);   // [256] first slot of each tile row
            int *coloff = rowoff + TILE;                 // [256] first slot of each tile column
            int *wsum = coloff + TILE;                   // [4]
            int *wsum2 = wsum + 4;                       // [4]
            unsigned long long *sbase = (unsigned long long *)(wsum2 + 4);
            int *wcnt = (int *)(sbase + 1);              // [4] kept pairs per wave
            uint64_t *rowpre = (uint64_t *)(epi + 16384 + 2048 + 64);  // [256] word prefixes
            uint64_t *colpre = rowpre + TILE;                           // [256]
            lds_u64 *lists = (lds_u64 *)(epi + EPI_FIXED);  // [4 waves][LIST_LDS]
            for (int k = etid; k < TILE * 8 * 2 / 4; k += CT_THREADS) ((u32x4 *)rowbits)[k] = (u32x4){0u, 0u, 0u, 0u};
            lds_sync();

            WaveList wl{lists + ewave * C::LIST_LDS,
                        (glb_u64 *)(spill + ((int64_t)blockIdx.x * 4 + ewave) * (WAVE_PAIRS - C::LIST_LDS)),
                        C::LIST_LDS, 0};
            const bool full = !(ec.diag || ec.gi0 + TILE > sd.N || ec.gj0 + TILE > sd.N);
            if (full)
                BNA_FOR_BLOCKS_T(epi_mark_inner, ec, rowbits, colbits, wl);
            else
                BNA_FOR_BLOCKS_T(epi_mark_edge, ec, rowbits, colbits, wl);
            if (elane == 0) wcnt[ewave] = wl.n;
            lds_sync();

            // The tile's entry range is reserved now, from the per-wave counts
            // (every kept pair is in one wave's list and sets one row bit), so
            // the atomic's round trip overlaps the popcounts and scans below.
            unsigned long long sb = 0;
            if (etid == 0) {
                const int tp = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
                if (tp) sb = gatomic_add(sd.entry_count, (unsigned long long)(2 * tp));
            }
            int rc = 0, cc = 0;
            uint64_t rpre = 0, cpre = 0;
#pragma unroll
            for (int w = 0; w < 8; w++) {
                rpre |= (uint64_t)rc << (8 * w);
                cpre |= (uint64_t)cc << (8 * w);
                rc += __popc(rowbits[etid * 8 + w]);
                cc += __popc(colbits[etid * 8 + w]);
            }
            rowpre[etid] = rpre;
            colpre[etid] = cpre;
            int total = 0, total2 = 0;
            const int ro = block_excl_scan(rc, wsum, total, elane, ewave);
            const int co = block_excl_scan(cc, wsum2, total2, elane, ewave);
            rowoff[etid] = ro;
            coloff[etid] = co;
            if (etid == 0) *sbase = sb;
            lds_sync();
            const int64_t base = (int64_t)*sbase;
            const bool fits = base + 2 * (int64_t)total <= sd.entry_cap;

            // fragment tables, [k][row] so that a tile's writes are contiguous
            const int N = sd.N;
            if (ec.gi0 + etid < N) {
                const int64_t f = (int64_t)(ic.J + 1) * N + ec.gi0 + etid;
                ((g_i32 *)sd.frag_orig)[f] = rc;
                ((g_i32 *)sd.frag_live)[f] = rc;
                ((g_i64 *)sd.frag_off)[f] = base + ro;
            }
            if (ec.gj0 + etid < N) {
                const int64_t f = (int64_t)ic.I * N + ec.gj0 + etid;
                ((g_i32 *)sd.frag_orig)[f] = cc;
                ((g_i32 *)sd.frag_live)[f] = cc;
                ((g_i64 *)sd.frag_off)[f] = base + total + co;
            }
            const SlotCtx sc{rowbits, colbits, rowoff, coloff, rowpre, colpre, base, total};
            if (wl.n > wl.cap) __builtin_amdgcn_s_waitcnt(0);  // this wave's spilled records are stored
            for (int k0 = 0; k0 < wl.n; k0 += 64) {
                const int k = k0 + elane;
                const bool act = k < wl.n;
                uint64_t raw = 0;
                if (act) raw = k < wl.cap ? (uint64_t)wl.lds[k] : __builtin_nontemporal_load(wl.spill + (k - wl.cap));
                const uint32_t lo32 = (uint32_t)raw;
                emit_pair(sd, ec, sc, fits, act, (int)(lo32 & 0xffff), (int)(lo32 >> 16),
                          __builtin_bit_cast(float, (uint32_t)(raw >> 32)));
            }
        }

        // ------------------------------------------------------------ boundary
        // Publish the claimed tile after next: the home range's answer, or, once
        // that range is drained, a claim from the others.
        if (etid == 0) {
            int64_t nn_claim = -1;
            if (nxt >= 0) {
                const int64_t lo = (int64_t)xcd * per_xcd;
                const int64_t cnt = n_items - lo < per_xcd ? n_items - lo : per_xcd;
                nn_claim = (int64_t)pre_t < cnt ? lo + (int64_t)pre_t : claim_tile(claim, n_items, per_xcd, xcd);
            }
            ctl[2] = nn_claim;
        }
        __syncthreads();
        const int64_t nn = uni64(ctl[2]);
        if constexpr (HIST == 1) {
            if (nxt < 0 || in.subj != ic.subj) {
                hist_flush((uint32_t *)epi, sd.hist, etid);
                __syncthreads();
            }
        }
        if (nxt < 0) break;
        cur = nxt;
        ic = items[cur];
        nxt = nn;
        in = nxt >= 0 ? items[nxt] : ic;
        g += nk;
    }
#undef BNA_DMA
#undef BNA_LDFRAG
#undef BNA_MFMA
#undef BNA_INTERLEAVE
}

// ------------------------------------------------------------ 8-wave form
// The same tile, stream and epilogue with eight waves, two per SIMD, each owning
// a 128 x 64 sub-tile (4 x 2 accumulator blocks, 128 AGPRs).  Waves 0-3 (group
// 0) and 4-7 (group 1) share SIMDs pairwise and alternate: while one group runs
// the MFMAs of a stage, the other reads its next stage's fragments from LDS and
// issues its quarter of a stage load, so the matrix pipe of each SIMD is fed by
// one wave while the other waits on memory and barriers.  Each group's waves
// run the phases M(s) (fragments of stage s into registers, load of stream
// position s + 3, waits) and C(s) (the MFMAs of stage s) with a workgroup
// barrier after every phase; group 1 starts one phase later, so its C(s)
// overlaps group 0's M(s + 1).  Group 0 loads the A side of every stage, group 1
// the B side.  Stage s is read by group 0 in half-period 2s and by group 1 in
// 2s + 1, so its slot is refilled (with position s + 4) from half-period 2s + 2
// on; each wave waits for its own part of stage s + 1 at the end of its M(s),
// before the barrier that precedes the first read of stage s + 1.
constexpr int CT8 = 512;
constexpr int WAVE_PAIRS8 = 128 * 64;

template <int PREC>
struct Corr8Cfg {
    using B = CorrCfg<PREC>;
    static constexpr int UNIT = B::UNIT, CH = B::CH, PANEL = B::PANEL, SLOT = B::SLOT, RING = B::RING;
    static constexpr int DPW = PANEL / 1024 / 4;  // 1 KB stage loads per wave and stage (a quarter of one side)
    static constexpr int LIST_RAW = (LDS_MAX - CTL_BYTES - RING - EPI_FIXED) / (8 * 8);
    static constexpr int LIST_LDS = (LIST_RAW > 2048 ? 2048 : LIST_RAW) & ~7;  // kept pairs per wave
};
static_assert(Corr8Cfg<2>::LIST_LDS >= 128, "8-wave epilogue lists do not fit beside the ring");
static_assert(HIST_BINS * 4 <= EPI_FIXED + 8 * 8 * Corr8Cfg<2>::LIST_LDS, "8-wave histogram region");

#define BNA_FOR_BLOCKS8_T(F, ...)                                                                  \
    do {                                                                                           \
        BNA_BLOCK_T(F, 0, 0, __VA_ARGS__); BNA_BLOCK_T(F, 0, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 1, 0, __VA_ARGS__); BNA_BLOCK_T(F, 1, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 2, 0, __VA_ARGS__); BNA_BLOCK_T(F, 2, 1, __VA_ARGS__);                      \
        BNA_BLOCK_T(F, 3, 0, __VA_ARGS__); BNA_BLOCK_T(F, 3, 1, __VA_ARGS__);                      \
    } while (0)
#define BNA_FOR_BLOCKS8(F, ...)                                                                    \
    do {                                                                                           \
        F(acc[0][0], 0, 0, __VA_ARGS__); F(acc[0][1], 0, 1, __VA_ARGS__);                          \
        F(acc[1][0], 1, 0, __VA_ARGS__); F(acc[1][1], 1, 1, __VA_ARGS__);                          \
        F(acc[2][0], 2, 0, __VA_ARGS__); F(acc[2][1], 2, 1, __VA_ARGS__);                          \
        F(acc[3][0], 3, 0, __VA_ARGS__); F(acc[3][1], 3, 1, __VA_ARGS__);                          \
    } while (0)

// Exclusive scan of one int per thread over 512 threads in two independent
// halves of four waves (threads 0-255 and 256-511).  lo_total: the first
// half's total, seen by every thread.
__device__ __forceinline__ int halves_excl_scan(int v, int *wsum, int &lo_total, int lane, int w) {
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) wsum[w] = inc;
    lds_sync();
    int pre = 0, t0 = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int s = wsum[k];
        if (k < 4) t0 += s;
        if ((k >> 2) == (w >> 2) && k < w) pre += s;
    }
    lo_total = t0;
    return pre + inc - v;
}

template <int PREC, int HIST>
__global__ __launch_bounds__(CT8, 1) void corr_tile8_kernel(const SubjDesc *__restrict__ subjs,
                                                            const TileItem *__restrict__ items, int64_t n_items,
                                                            unsigned int *__restrict__ claim,
                                                            uint64_t *__restrict__ spill)
{
    static_assert(PREC == 1 || PREC == 2, "the 8-wave form runs the bf16 precisions");
    static_assert(HIST == 0 || HIST == 1, "the 8-wave form runs the emission and histogram passes");
    using C = Corr8Cfg<PREC>;
    extern __shared__ __align__(1024) unsigned char smem[];
    unsigned char *const epi = smem + C::RING;
    int64_t *const ctl = (int64_t *)(smem + LDS_MAX - CTL_BYTES);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int swave = __builtin_amdgcn_readfirstlane(wave);
    const int grp = swave >> 2, quarter = swave & 3;  // group, and quarter of the group's side it loads
    const int wr0 = (wave & 1) * 128;
    const int wc0 = (wave >> 1) * 64;
    const int lrow = lane & 31;
    const int half = lane >> 5;
    const int xcd = blockIdx.x & 7;
    const int64_t per_xcd = (n_items + 7) / 8;

    if (tid == 0) {
        const int64_t t0 = claim_tile(claim, n_items, per_xcd, xcd);
        ctl[0] = t0;
        ctl[1] = t0 >= 0 ? claim_tile(claim, n_items, per_xcd, xcd) : -1;
    }
    if constexpr (HIST == 1) {
        for (int k = tid; k < HIST_BINS; k += CT8) ((uint32_t *)epi)[k] = 0;
    }
    __syncthreads();
    int64_t cur = uni64(ctl[0]);
    int64_t nxt = uni64(ctl[1]);
    if (cur < 0) return;
    TileItem ic = items[cur];
    TileItem in = nxt >= 0 ? items[nxt] : ic;

    // Stage loads: wave (grp, quarter) fills rows [quarter * 64, quarter * 64 + 64)
    // of side grp (0: A, 1: B) with DPW 1 KB instructions, addressed as in the
    // 4-wave form.
    const uint32_t dsrc = (uint32_t)((quarter * (TILE / 4) + lane / C::CH) * C::UNIT +
                                     (((lane % C::CH) ^ swz<C::CH>(lane / C::CH)) << 4));
    int g = 0;  // stream position of the current tile's stage 0; stage s lives in slot (g + s) % NSLOT

// Source and LDS destination of this wave's part of stream position S.
#define BNA8_DMA_ADDR(S, P, D)                                                                     \
    do {                                                                                           \
        const int s_ = (S);                                                                        \
        const bool own_ = s_ < nk;                                                                 \
        const int64_t o_ = own_ ? s_ * ic.kstride : (s_ - nk) * xK;                                \
        P = (own_ ? (grp ? cB : cA) : (grp ? xB : xA)) + o_ + dsrc;                                \
        D = smem + ((g + s_) & (NSLOT - 1)) * C::SLOT + grp * C::PANEL + quarter * C::DPW * 1024;  \
    } while (0)
#define BNA8_DMA(S)                                                                                \
    do {                                                                                           \
        const unsigned char *p_;                                                                   \
        unsigned char *d_;                                                                         \
        BNA8_DMA_ADDR(S, p_, d_);                                                                  \
        dma_side<C::DPW>(p_, d_);                                                                  \
    } while (0)

    u32x4 fa[4][2], fb[2][2];
    const int c0 = half, c1 = 2 + half;  // logical chunks: hi of lane half h, lo of lane half h
    const int sw = swz<C::CH>(lrow);
    const uint32_t fA0 = (uint32_t)((wr0 + lrow) * C::UNIT + ((c0 ^ sw) << 4));
    const uint32_t fA1 = (uint32_t)((wr0 + lrow) * C::UNIT + ((c1 ^ sw) << 4));
    const uint32_t fB0 = (uint32_t)(C::PANEL + (wc0 + lrow) * C::UNIT + ((c0 ^ sw) << 4));
    const uint32_t fB1 = (uint32_t)(C::PANEL + (wc0 + lrow) * C::UNIT + ((c1 ^ sw) << 4));
#define BNA8_LDFRAG(STAGE_IDX)                                                                     \
    do {                                                                                           \
        const unsigned char *slot_ = smem + ((g + (STAGE_IDX)) & (NSLOT - 1)) * C::SLOT;           \
        _Pragma("unroll") for (int blk = 0; blk < 4; blk++) {                                      \
            fa[blk][0] = *(const u32x4 *)(slot_ + fA0 + blk * 32 * C::UNIT);                       \
            if constexpr (PREC == 2) fa[blk][1] = *(const u32x4 *)(slot_ + fA1 + blk * 32 * C::UNIT); \
        }                                                                                          \
        _Pragma("unroll") for (int blk = 0; blk < 2; blk++) {                                      \
            fb[blk][0] = *(const u32x4 *)(slot_ + fB0 + blk * 32 * C::UNIT);                       \
            if constexpr (PREC == 2) fb[blk][1] = *(const u32x4 *)(slot_ + fB1 + blk * 32 * C::UNIT); \
        }                                                                                          \
    } while (0)
#define BNA8_MFMA()                                                                                \
    do {                                                                                           \
        _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                           \
            _Pragma("unroll") for (int cb = 0; cb < 2; cb++)                                       \
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                             \
                    __builtin_bit_cast(bf16x8_t, fa[rb][0]), __builtin_bit_cast(bf16x8_t, fb[cb][0]), \
                    acc[rb][cb], 0, 0, 0);                                                         \
        if constexpr (PREC == 2) {                                                                 \
            _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                       \
                _Pragma("unroll") for (int cb = 0; cb < 2; cb++)                                   \
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                         \
                        __builtin_bit_cast(bf16x8_t, fa[rb][0]), __builtin_bit_cast(bf16x8_t, fb[cb][1]), \
                        acc[rb][cb], 0, 0, 0);                                                     \
            _Pragma("unroll") for (int rb = 0; rb < 4; rb++)                                       \
                _Pragma("unroll") for (int cb = 0; cb < 2; cb++)                                   \
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(                         \
                        __builtin_bit_cast(bf16x8_t, fa[rb][1]), __builtin_bit_cast(bf16x8_t, fb[cb][0]), \
                        acc[rb][cb], 0, 0, 0);                                                     \
        }                                                                                          \
    } while (0)

    // Prologue of the first tile: stream positions 0, 1, 2.
    {
        const int nk = ic.nk;
        const unsigned char *cA = (const unsigned char *)ic.panel + (int64_t)ic.I * C::PANEL;
        const unsigned char *cB = (const unsigned char *)ic.panel + (int64_t)ic.J * C::PANEL;
        const unsigned char *xA = cA, *xB = cB;
        const int64_t xK = 0;
        BNA8_DMA(0);
        BNA8_DMA(1);
        BNA8_DMA(2);
    }

    for (;;) {
        const int nk = ic.nk;
        const unsigned char *cA = (const unsigned char *)ic.panel + (int64_t)ic.I * C::PANEL;
        const unsigned char *cB = (const unsigned char *)ic.panel + (int64_t)ic.J * C::PANEL;
        const unsigned char *xA = nxt >= 0 ? (const unsigned char *)in.panel + (int64_t)in.I * C::PANEL
                                           : cA + (int64_t)(nk - 1) * ic.kstride;
        const unsigned char *xB = nxt >= 0 ? (const unsigned char *)in.panel + (int64_t)in.J * C::PANEL
                                           : cB + (int64_t)(nk - 1) * ic.kstride;
        const int64_t xK = nxt >= 0 ? in.kstride : 0;
        const SubjDesc sd = subjs[ic.subj];

        f32x16_t acc[4][2];
#pragma unroll
        for (int rb = 0; rb < 4; rb++)
#pragma unroll
            for (int cb = 0; cb < 2; cb++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[rb][cb][e] = 0.0f;

        // Everything in flight is waited for first (the previous tile's epilogue
        // stores share the counter with the stage loads), then a barrier makes
        // stages 0-2 visible to every wave.
        BNA_WAIT_VMCNT(0);
        unsigned int pre_t = 0;
        if (tid == 0 && nxt >= 0) pre_t = atomicAdd(&claim[xcd], 1u);
        __builtin_amdgcn_s_barrier();
        if (grp == 1) __builtin_amdgcn_s_barrier();  // group 1 runs one phase behind
        __builtin_amdgcn_sched_barrier(0);
        for (int kt = 0; kt < nk; kt++) {
            // M(kt): fragments of stage kt; then this wave's part of stage kt + 1
            // (loaded during its C(kt - 2)) has landed, position kt + 2 may stay
            // in flight, and its fragment reads are complete.
            BNA8_LDFRAG(kt);
            const unsigned char *dp;
            unsigned char *dd;
            BNA8_DMA_ADDR(kt + 3, dp, dd);
            BNA_WAIT_VMCNT(C::DPW);
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
            // C(kt): the MFMAs of stage kt, with this wave's part of stream
            // position kt + 3 issued between them
            dma_side<C::DPW>(dp, dd);
            BNA8_MFMA();
#pragma unroll
            for (int i_ = 0; i_ < C::DPW; i_++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, (PREC == 2 ? 24 : 8) - C::DPW, 0);
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_sched_barrier(0);
        }
        if (grp == 0) __builtin_amdgcn_s_barrier();  // group 1's last C phase
        __builtin_amdgcn_sched_barrier(0);

        // ------------------------------------------------------------ epilogue
        int etid = tid;
        asm volatile("" : "+v"(etid));
        const int elane = etid & 63, ewave = etid >> 6;
        EpiCtx ec;
        ec.N = sd.N;
        ec.gi0 = ic.I * TILE;
        ec.gj0 = ic.J * TILE;
        ec.wr0 = (ewave & 1) * 128;
        ec.wc0 = (ewave >> 1) * 64;
        ec.half = elane >> 5;
        ec.lrow = elane & 31;
        ec.lane = elane;
        ec.diag = (ic.I == ic.J);
        ec.keep = sd.keep_thr;

        if constexpr (HIST == 1) {
            BNA_FOR_BLOCKS8(epi_hist, ec, (uint32_t *)epi);
        } else {
            uint32_t *rowbits = (uint32_t *)epi;         // [256][8] tile-row bitmaps
            uint32_t *colbits = rowbits + TILE * 8;      // [256][8] tile-column bitmaps
            int *rowoff = (int *)(colbits + TILE * 8);   // [256] first slot of each tile row
            int *coloff = rowoff + TILE;                 // [256] first slot of each tile column
            int *wsum = coloff + TILE;                   // [8]
            int *wcnt = wsum + 8;                        // [8] kept pairs per wave
            uint64_t *rowpre = (uint64_t *)(epi + 16384 + 2048 + 64);  // [256] word prefixes
            uint64_t *colpre = rowpre + TILE;                           // [256]
            lds_u64 *lists = (lds_u64 *)(epi + EPI_FIXED);  // [8 waves][LIST_LDS]
            for (int k = etid; k < TILE * 8 * 2 / 4; k += CT8) ((u32x4 *)rowbits)[k] = (u32x4){0u, 0u, 0u, 0u};
            lds_sync();

            WaveList wl{lists + ewave * C::LIST_LDS,
                        (glb_u64 *)(spill + ((int64_t)blockIdx.x * 8 + ewave) * (WAVE_PAIRS8 - C::LIST_LDS)),
                        C::LIST_LDS, 0};
            const bool full = !(ec.diag || ec.gi0 + TILE > sd.N || ec.gj0 + TILE > sd.N);
            if (full)
                BNA_FOR_BLOCKS8_T(epi_mark_inner, ec, rowbits, colbits, wl);
            else
                BNA_FOR_BLOCKS8_T(epi_mark_edge, ec, rowbits, colbits, wl);
            if (elane == 0) wcnt[ewave] = wl.n;
            lds_sync();

            unsigned long long sb = 0;
            if (etid == 0) {
                int tp = 0;
#pragma unroll
                for (int w = 0; w < 8; w++) tp += wcnt[w];
                if (tp) sb = gatomic_add(sd.entry_count, (unsigned long long)(2 * tp));
            }
            // threads 0-255: tile row etid; threads 256-511: tile column etid - 256
            const bool is_row = etid < TILE;
            const int rr = etid & (TILE - 1);
            const uint32_t *bits = is_row ? rowbits : colbits;
            int cnt = 0;
            uint64_t pre = 0;
#pragma unroll
            for (int w = 0; w < 8; w++) {
                pre |= (uint64_t)cnt << (8 * w);
                cnt += __popc(bits[rr * 8 + w]);
            }
            (is_row ? rowpre : colpre)[rr] = pre;
            int total = 0;
            const int o = halves_excl_scan(cnt, wsum, total, elane, ewave);
            (is_row ? rowoff : coloff)[rr] = o;
            if (etid == 0) ctl[3] = (int64_t)sb;
            lds_sync();
            const int64_t base = ctl[3];
            const bool fits = base + 2 * (int64_t)total <= sd.entry_cap;

            const int N = sd.N;
            if (is_row) {
                if (ec.gi0 + rr < N) {
                    const int64_t f = (int64_t)(ic.J + 1) * N + ec.gi0 + rr;
                    ((g_i32 *)sd.frag_orig)[f] = cnt;
                    ((g_i32 *)sd.frag_live)[f] = cnt;
                    ((g_i64 *)sd.frag_off)[f] = base + o;
                }
            } else if (ec.gj0 + rr < N) {
                const int64_t f = (int64_t)ic.I * N + ec.gj0 + rr;
                ((g_i32 *)sd.frag_orig)[f] = cnt;
                ((g_i32 *)sd.frag_live)[f] = cnt;
                ((g_i64 *)sd.frag_off)[f] = base + total + o;
            }
            const SlotCtx sc{rowbits, colbits, rowoff, coloff, rowpre, colpre, base, total};
            if (wl.n > wl.cap) __builtin_amdgcn_s_waitcnt(0);  // this wave's spilled records are stored
            for (int k0 = 0; k0 < wl.n; k0 += 64) {
                const int k = k0 + elane;
                const bool act = k < wl.n;
                uint64_t raw = 0;
                if (act) raw = k < wl.cap ? (uint64_t)wl.lds[k] : __builtin_nontemporal_load(wl.spill + (k - wl.cap));
                const uint32_t lo32 = (uint32_t)raw;
                emit_pair(sd, ec, sc, fits, act, (int)(lo32 & 0xffff), (int)(lo32 >> 16),
                          __builtin_bit_cast(float, (uint32_t)(raw >> 32)));
            }
        }

        // ------------------------------------------------------------ boundary
        if (etid == 0) {
            int64_t nn_claim = -1;
            if (nxt >= 0) {
                const int64_t lo = (int64_t)xcd * per_xcd;
                const int64_t cnt = n_items - lo < per_xcd ? n_items - lo : per_xcd;
                nn_claim = (int64_t)pre_t < cnt ? lo + (int64_t)pre_t : claim_tile(claim, n_items, per_xcd, xcd);
            }
            ctl[2] = nn_claim;
        }
        __syncthreads();
        const int64_t nn = uni64(ctl[2]);
        if constexpr (HIST == 1) {
            if (nxt < 0 || in.subj != ic.subj) {
                for (int k = etid; k < HIST_BINS; k += CT8) {
                    const uint32_t h = ((uint32_t *)epi)[k];
                    if (h) {
                        gatomic_add(sd.hist + k, (unsigned long long)h);
                        ((uint32_t *)epi)[k] = 0;
                    }
                }
                __syncthreads();
            }
        }
        if (nxt < 0) break;
        cur = nxt;
        ic = items[cur];
        nxt = nn;
        in = nxt >= 0 ? items[nxt] : ic;
        g += nk;
    }
#undef BNA8_DMA
#undef BNA8_DMA_ADDR
#undef BNA8_LDFRAG
#undef BNA8_MFMA
}

// Clears the per-range claim counters of one launch.
__global__ void claim_reset_kernel(unsigned int *claim)
{
    if (threadIdx.x < 8) claim[threadIdx.x] = 0u;
}

template <int PREC, int HIST, bool EIGHT>
static hipError_t launch_one(const SubjDesc *subj_dev, const TileItem *items, int64_t n_items, unsigned int *claim,
                             uint64_t *spill, int grid_max, hipStream_t st)
{
    static std::atomic<int> attr_dev{-1};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const void *fn;
    if constexpr (EIGHT)
        fn = (const void *)corr_tile8_kernel<PREC, HIST>;
    else
        fn = (const void *)corr_tile_kernel<PREC, HIST>;
    if (attr_dev.load() != dev) {
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, LDS_MAX);
        if (e != hipSuccess) return e;
        attr_dev.store(dev);
    }
    if (n_items <= 0) return hipSuccess;
    if (grid_max < 8 || grid_max % 8) return hipErrorInvalidValue;
    const int64_t grid = std::min<int64_t>(grid_max, (n_items + 7) / 8 * 8);
    hipLaunchKernelGGL(claim_reset_kernel, dim3(1), dim3(64), 0, st, claim);
    if constexpr (EIGHT)
        hipLaunchKernelGGL((corr_tile8_kernel<PREC, HIST>), dim3((unsigned)grid), dim3(CT8), LDS_MAX, st, subj_dev,
                           items, n_items, claim, spill);
    else
        hipLaunchKernelGGL((corr_tile_kernel<PREC, HIST>), dim3((unsigned)grid), dim3(CT_THREADS), LDS_MAX, st,
                           subj_dev, items, n_items, claim, spill);
    return hipGetLastError();
}

// The bf16 precisions' emission and histogram passes run the 8-wave form
// unless BNA_TILE_WAVES=4 is set in the environment (read once per process).
static bool eight_waves()
{
    static const bool on = [] {
        const char *v = getenv("BNA_TILE_WAVES");
        return !(v && atoi(v) == 4);
    }();
    return on;
}

hipError_t launch_corr(int prec, int hist_mode, const SubjDesc *subj_dev, const TileItem *items,
                       int64_t n_items, unsigned int *claim, uint64_t *spill, int grid_max, hipStream_t st)
{
    if (hist_mode < 0 || hist_mode > 2) return hipErrorInvalidValue;
    const bool e8 = eight_waves() && prec != 0 && hist_mode != 2;
    switch (prec * 3 + hist_mode) {
    case 0: return launch_one<0, 0, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 1: return launch_one<0, 1, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 2: return launch_one<0, 2, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 3: return e8 ? launch_one<1, 0, true>(subj_dev, items, n_items, claim, spill, grid_max, st)
                      : launch_one<1, 0, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 4: return e8 ? launch_one<1, 1, true>(subj_dev, items, n_items, claim, spill, grid_max, st)
                      : launch_one<1, 1, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 5: return launch_one<1, 2, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 6: return e8 ? launch_one<2, 0, true>(subj_dev, items, n_items, claim, spill, grid_max, st)
                      : launch_one<2, 0, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 7: return e8 ? launch_one<2, 1, true>(subj_dev, items, n_items, claim, spill, grid_max, st)
                      : launch_one<2, 1, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    case 8: return launch_one<2, 2, false>(subj_dev, items, n_items, claim, spill, grid_max, st);
    default: return hipErrorInvalidValue;
    }
}

// Records of the per-wave spill area for a launch of at most grid_max workgroups.
int64_t corr_spill_records(int prec, int grid_max)
{
    const int cap = prec == 0 ? CorrCfg<0>::LIST_LDS : prec == 1 ? CorrCfg<1>::LIST_LDS : CorrCfg<2>::LIST_LDS;
    const int cap8 = prec == 1 ? Corr8Cfg<1>::LIST_LDS : Corr8Cfg<2>::LIST_LDS;
    return (int64_t)grid_max * std::max<int64_t>(4 * (WAVE_PAIRS - cap), 8 * (WAVE_PAIRS8 - cap8));
}

}  // namespace bna
