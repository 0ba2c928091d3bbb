// common.hpp -- shared definitions of libbna's kernels and host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace bna {

// Correlation tile edge (voxels), the K padding of the operand panels and the
// shortest padded length (the tile kernel streams at least four 16-point stages).
constexpr int TILE = 256;
constexpr int KPAD = 16;
constexpr int LPAD_MIN = 64;

// One entry of the per-tile compacted edge lists: the neighbour's voxel index
// and its r.  col < 0 marks an entry the FP64 recheck rejected.
struct Entry {
    int32_t col;
    float r;
};

// A pair inside the error band, queued for the FP64 recheck.
struct Candidate {
    int32_t i, j;       // i < j, global voxel indices
    int64_t slot_row;   // entry index of (i -> j) in the row-view list
    int64_t slot_col;   // entry index of (j -> i) in the column-view list
    float r;            // the MFMA value
    int32_t pad;
};

// Tile work item: the subject's operand panel and its stride between 16-point
// blocks (N_pad * panel unit), the subject index, tile coordinates (I <= J) and
// the number of 16-point stages (L_pad / 16, at least 4).
struct TileItem {
    const void *panel;
    int64_t kstride;
    int32_t subj;
    int32_t I, J;
    int32_t nk;
};

// Per-subject description used by the tile kernels (device memory).
struct SubjDesc {
    const void *a0;      // k-blocked operand panel (see panel_unit)
    const void *a1;      // unused
    int32_t N, N_pad, L_pad, nb;
    float keep_thr;      // entry kept iff r > keep_thr
    float sure_thr;      // kept entry is sure iff r > sure_thr, else a candidate
    int32_t *frag_orig;  // (nb + 1) x N, fragment-major: [k * N + row]
    int32_t *frag_live;  // (nb + 1) x N
    int64_t *frag_off;   // (nb + 1) x N
    Entry *entries;      // compacted lists
    unsigned long long *entry_count; // running total (may exceed capacity)
    int64_t entry_cap;
    Candidate *cands;
    unsigned long long *cand_count;
    int64_t cand_cap;
    unsigned long long *hist; // histogram pass: HIST_BINS counters
    float *dense;        // accumulation pass: packed upper triangle, M1 floats
    int32_t fisher;      // accumulation pass: add Fisher z instead of r
    uint8_t *mirror;     // per row-view entry slot: position of the pair's other entry within
                         // its column-view fragment (null: not recorded)
    uint32_t *flags;     // accumulation pass, Fisher sums: bit p set when some subject's |r| of
                         // packed pair p exceeds flag_r (null: not recorded)
    float flag_r;
};

constexpr int HIST_BINS = 4096;  // over [-1, 1], clamped at both ends

// Radix-select state for the density threshold (device memory).
struct SelState {
    unsigned long long key_hi, key_lo;  // threshold key fixed so far (r part, pair part)
    long long rank;                     // rank of the threshold key among keys matching it
    int all;                            // 1: the threshold key is final
    int err;                            // nonzero: counts inconsistent, nothing selected
    unsigned int done, pad;             // workgroups finished in the current pass
    unsigned long long kmax;            // largest kept r key (gives r*)
    long long n_sure;                   // sure edges (pairs emitted above the bracket)
    unsigned long long hist[256];
};

// bf16 round-to-nearest-even from float bits; NaN stays a quiet NaN.
__host__ __device__ inline uint16_t f32_to_bf16_rn(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// ------------------------------------------------------------ launchers
// Bytes per (16-point block, voxel row) record of the k-blocked operand panel:
// fp32 z[16] = 64, bf16 hi[16] = 32, bf16x3 hi[16] lo[16] = 64.
__host__ __device__ inline int panel_unit(int prec) { return prec == 1 ? 32 : 64; }

// z-normalisation (kernel a): mean, norm, validity and the k-blocked MFMA operand
// panel; for time-major input (layout 0) also xv, the series voxel-major with row
// stride L_pad.
hipError_t launch_znorm(const float *x, int layout, int N, int L, int N_pad, int L_pad, double *mu,
                        double *sd, float *xv, uint8_t *valid, int prec, unsigned char *panel, hipStream_t st);
// z64 [N][L] from a voxel-major series with row stride ld.
hipError_t launch_zmat(const float *xs, int64_t ld, int N, int L, const double *mu, const double *sd, double *z64,
                       hipStream_t st);

// Correlation + threshold + compaction over a tile table (kernel b + first half of c),
// as one persistent launch of at most grid_max workgroups.  prec: 0 FP32, 1 BF16,
// 2 BF16X3.  hist_mode: 1 = histogram epilogue, 2 = add every pair's r (or
// Fisher z) to the subject's packed dense sums.  claim: 8 device counters (reset
// by the launch).  spill: corr_spill_records(prec, grid_max) 8-byte records.
// grid_max: a positive multiple of 8.
hipError_t launch_corr(int prec, int hist_mode, const SubjDesc *subj_dev, const TileItem *items,
                       int64_t n_items, unsigned int *claim, uint64_t *spill, int grid_max, hipStream_t st);
int64_t corr_spill_records(int prec, int grid_max);

// Group averaging over a packed upper triangle of M1 floats (k_group.hip).
// Divide every sum by the subject count S in float and, for Fisher sums, map
// back with the inverse transform.
hipError_t launch_group_finish(float *acc, int64_t M1, int S, int fisher, hipStream_t st);
// One subject of an exact group: its series voxel-major with row stride ld (the
// normalisation's xv, or the caller's voxel-major x), mean and norm per voxel.
struct GroupSubj {
    const float *xs;
    int64_t ld;
    const double *mu, *sd;
    int32_t L, pad;
};
// Pairs (i < j) of the packed triangle whose average lies within w of t, or whose
// bit is set in flags (may be null), appended to pairs[] (count[0] counts every
// match, also past cap).
hipError_t launch_group_cands(const float *acc, int N, float t, float w, const uint32_t *flags, int2 *pairs,
                              unsigned long long *count, int64_t cap, hipStream_t st);
// For each listed pair: the average the reference's arithmetic gives from the
// subjects' r64 (each taken as float, or its Fisher z, added in subject order in
// float when it lies in [-1, 1]; divided by S in float; mapped back for Fisher
// sums), written into acc at the pair's packed index.
hipError_t launch_group_exact(const int2 *pairs, const unsigned long long *count, int64_t cap,
                              const GroupSubj *subjs, int S, int fisher, float *acc, int N, hipStream_t st);
// The entries (i, j) with j > i of a device CSR, appended to pairs[].
hipError_t launch_csr_upper_pairs(const int32_t *R, const int32_t *C, int N, int2 *pairs,
                                  unsigned long long *count, int64_t cap, hipStream_t st);
// V[e] = acc at the packed index of entry e's pair.
hipError_t launch_group_refill(const int32_t *R, const int32_t *C, const float *acc, int N, float *V,
                               hipStream_t st);
// 256-bin histogram of one radix digit of the order-preserving keys of the
// values whose higher digits equal `prefix` under `mask` (added to hist).
hipError_t launch_group_digits(const float *acc, int64_t M1, uint32_t prefix, uint32_t mask, int shift,
                               unsigned long long *hist, hipStream_t st);
// Per packed row i: how many of its values equal `val` (ties of a density threshold).
hipError_t launch_group_ties(const float *acc, int N, float val, int64_t *tie_rows, hipStream_t st);
// Full N x N keep matrix (bytes, both triangles, zero diagonal, row stride ld =
// N rounded up to 64, padding columns 0): density mode (dens = 1) keeps v > thr
// and the ties v == thr at packed index <= p_cut; r mode keeps v > thr.
// flag[0] gets 1 if any value exceeds 1.
hipError_t launch_group_keep(const float *acc, int N, int64_t ld, float thr, int dens, int64_t p_cut, uint8_t *keep,
                             unsigned int *flag, hipStream_t st);
// Row lengths of the keep matrix, then the CSR rows (columns ascending, values
// from the packed averages when V is not null).
hipError_t launch_group_rowcount(const uint8_t *keep, int N, int64_t ld, int64_t *cnt, hipStream_t st);
hipError_t launch_group_fill(const uint8_t *keep, int64_t ld, const float *acc, int N, const int32_t *R, int32_t *C,
                             float *V, hipStream_t st);

// FP64 recheck of candidates.
// Slots at or past entry_cap (overflowed lists) are skipped.
// z is recomputed from the voxel-major series xs (row stride ld), mu and sd.
// mode 0: decide r64 > thr; mode 1: write r64 to r64_out and to both entry slots
// (density selection decides later); mode 2: as 1 with the MFMA r taken as is.
hipError_t launch_recheck(const float *xs, int64_t ld, const double *mu, const double *sd, int N, int L, int nb,
                          const Candidate *cands,
                          const unsigned long long *cand_count, int64_t cand_cap, double thr,
                          int32_t *frag_live, Entry *entries, int64_t entry_cap, unsigned long long *stats,
                          int mode, double *r64_out, hipStream_t st);
// Exact-k selection among rechecked candidates (ctr: [0] entry slots, [1] candidates).
hipError_t launch_select(SelState *st, const unsigned long long *ctr, const Candidate *cands, const double *r64,
                         int64_t cand_cap, int64_t k, int N, int32_t *frag_live, Entry *entries, int64_t entry_cap,
                         hipStream_t stream);

// Row lengths, exclusive scan into R, placement into C/V.
// Fragment tables [k][row] -> [row][k], plus row lengths (sums of live counts).
hipError_t launch_frag_rows(const int32_t *orig_km, const int32_t *live_km, const int64_t *off_km, int N, int nb,
                            int32_t *orig_rm, int32_t *live_rm, int64_t *off_rm, int64_t *rowlen, hipStream_t st);
// bsum: scan_blocks(N) + 1 int64 of scratch.
hipError_t launch_scan(const int64_t *rowlen, int N, int64_t *R64, int32_t *R32, int64_t *bsum, hipStream_t st);
int scan_blocks(int N);
// Recomputes every value of a weighted CSR in double from the voxel-major series.
// Rows [row0, row1) of the CSR; nnz is the whole CSR's entry count (sizes the grid).
hipError_t launch_zero_u64(void *p, int64_t n, hipStream_t st);

// Shortest paths of a binary graph (k_paths.hip).  Multi-source BFS over 256
// sources [s0, s0 + ns) at a time: bit planes of 4 x 64-bit words per vertex.
hipError_t launch_msbfs_init(int N, int s0, int ns, void *visited, void *front, hipStream_t st);
// One level of the batch's searches; count += the pairs first reached at it.
hipError_t launch_msbfs_level(const int32_t *R, const int32_t *C, int N, const void *front, void *next,
                              void *visited, int ns, unsigned long long *count, hipStream_t st);
// Brandes betweenness of sources [s0, s0 + ns): per-source state arrays of ns x N
// (lvl: ns x (N + 2)), out ns x N; the deltas are added into bc in source order.
hipError_t launch_bc_batch(const int32_t *R, const int32_t *C, int N, int s0, int ns, int32_t *dist,
                           int32_t *queue, int32_t *lvl, double *sigma, double *delta, double *out, double *bc,
                           hipStream_t st);
// All-pairs shortest paths of a weighted graph (k_fw.hip): blocked
// Floyd-Warshall on an Np x Np float cost matrix, Np a multiple of FW_B.  The
// block size fixes which intermediate values each phase uses, hence the float
// distances.
constexpr int FW_B = 128;
// D = +inf, 0 on the diagonal, 1.0f / w (1 without weights) per CSR entry;
// bad |= 1 for a negative or NaN weight, |= 2 for a column out of range.
hipError_t launch_fw_init(const int32_t *R, const int32_t *C, const float *V, int N, int Np, float *D, unsigned *bad,
                          hipStream_t st);
// The three phases of block K.
hipError_t launch_fw_block(float *D, int Np, int K, hipStream_t st);
// eff[u] (nodal efficiency, float sum of 1.0f / d in column order) and
// rowsum[u] (the same terms summed in double) of rows u < N.
hipError_t launch_fw_rows(const float *D, int Np, int N, float *eff, double *rowsum, hipStream_t st);
hipError_t launch_narrow_rows(const int64_t *R64, int64_t n, int32_t *R32, hipStream_t st);
bool refine_lists_supported(const float *xs, int64_t ld, int L);
hipError_t launch_refine_lists(Entry *entries, const uint8_t *mirror, const int32_t *orig_rm, const int64_t *off_rm,
                               const int64_t *off_km, const float *xs, int64_t ld, const double *mu, const double *sd,
                               int N, int L, int nb, int row0, int row1, hipStream_t st);
hipError_t launch_refine(const int32_t *R, const int32_t *C, float *V, const float *xs, int64_t ld,
                         const double *mu, const double *sd, int N, int L, int row0, int row1, int64_t nnz,
                         hipStream_t st);
// frag_* here are the row-major tables.
hipError_t launch_place(const int32_t *frag_orig, const int32_t *frag_live, const int64_t *frag_off,
                        const Entry *entries, const int64_t *R64, int N, int nb, int32_t *C,
                        float *V, hipStream_t st);

// Degree and eigenvector centrality (kernel d).
hipError_t launch_degree(const int32_t *R, const float *V, int N, int64_t nnz, float *deg, hipStream_t st);
// How the power iteration's SpMV runs: segment-blocked with x staged in LDS
// (blocked), or the plain L2-gather kernel.  blocked needs N >= 4096, 16 or
// more entries per row on average and a layout that fits LDS: G workgroups of
// RB rows, S segments of SEG columns, lds bytes of dynamic LDS; items = G S RB
// (block, segment, row) slices and at most quads 4-entry quads.
struct EcPlan {
    bool blocked = false;
    int G = 0, RB = 0, S = 0, SEG = 0;
    size_t lds = 0;
    int64_t items = 0, quads = 0;
};
EcPlan ec_plan(int N, int64_t nnz, int ncu);
// Workspace of the blocked plan: qc N S ints, qoff items + 1 ints, btot G
// ints, cq 4 quads uint16, vq 4 quads floats (weighted graphs only).
struct EcBlocked {
    int32_t *qc = nullptr, *qoff = nullptr, *btot = nullptr;
    uint16_t *cq = nullptr;
    float *vq = nullptr;
};
// Whole power iteration (see k_graph.hip).  output: 0 last A*y, 1 last y.
// state_host: two EcState slots (2 * EC_STATE_MAX_BYTES bytes), pinned for the
// batches to overlap; ev: two events for the batches in flight.
// use_graph: replay batches of 16 iterations from one captured HIP graph (the
// plain plan only: the blocked plan's launches are long enough on their own).
hipError_t run_eigcent(const int32_t *R, const int32_t *C, const float *V, int N, const EcPlan &plan,
                       const EcBlocked &bl, double tol, int max_iter, int output, double *xa, double *xb, double *part,
                       void *state_dev, void *state_host, double *out, double *lambda, int *iters, bool use_graph,
                       const hipEvent_t ev[2], hipStream_t st);
size_t ec_state_bytes();
int ec_nparts(int N);
// Upper bounds of the iteration's scratch, for callers that size it up front:
// the state struct's bytes, and ec_nparts(N) for every N up to 4194304 (the
// blocked plan uses G <= CU count partials, fewer than either).
constexpr size_t EC_STATE_MAX_BYTES = 256;
constexpr int EC_NPARTS_MAX = 16384;

// Newman leading-eigenvector modularity (k_modularity.hip).  Device buffers:
// AD, cnt, res N ints; SR N + 1 ints; SC nnz ints; sumBG, y[0..2], yg[0..1]
// N doubles; part 4 NM_BLOCKS_MAX doubles; state nm_state_bytes(); same one
// u64; groups, gout, side N entries each.  state_host: pinned, two states; ev: two events.  R_host: the CSR's row
// offsets on the host.
constexpr int NM_BLOCKS_MAX = 8192;
struct NmGroup {     // a group's rows [a0, a1) of the batch, and its degree sum
    int32_t a0, a1;
    double inner;
};
struct NmGroupOut {  // a small group's eigenvalue, split counts, iterations
    double lambda;
    long long subP, subN;
    int32_t iters, pad;
};
struct NmBuffers {
    int32_t *AD = nullptr, *cnt = nullptr, *res = nullptr, *SR = nullptr, *SC = nullptr;
    double *sumBG = nullptr, *y[3] = {}, *yg[2] = {}, *part = nullptr;
    void *state = nullptr, *state_host = nullptr;
    unsigned long long *same = nullptr;
    NmGroup *groups = nullptr;     // N entries
    NmGroupOut *gout = nullptr;    // N entries
    int8_t *side = nullptr;        // N entries
    hipEvent_t ev[2] = {};
    std::vector<int32_t> AD_host;
    const int32_t *R_host = nullptr;
};
size_t nm_state_bytes();
// The whole partition: labels (host, N) numbered from 1, Q, module count,
// power iterations over all rounds, rounds.  V is not used (the reference
// tool reads binary graphs).
hipError_t run_modularity(const int32_t *R, const int32_t *C, int N, int64_t nnz, int max_iter, double eps,
                          NmBuffers &b, int32_t *labels, double *Q, int *modules, int64_t *iters_total,
                          int *rounds_out, hipStream_t st);

// Weighted clustering coefficient (k_cp.hip).  Columns must ascend within each
// row.  w: nnz floats of scratch (the cube-rooted weights); key: one u64 of
// scratch; cp: N floats.  V == NULL: binary graph.  N at most CP_MAX_N (the
// row kernel's LDS bitmap takes 8 bytes per 32 voxels).
constexpr int CP_MAX_N = 524288;
hipError_t launch_cp(const int32_t *R, const int32_t *C, const float *V, int N, int64_t nnz, float *w,
                     unsigned long long *key, float *cp, hipStream_t st);

}  // namespace bna
