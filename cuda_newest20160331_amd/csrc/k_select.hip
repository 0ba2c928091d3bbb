// k_select.hip -- exact-k edge selection for the density threshold.
//
// The reference (FormFullGraph_s, W/CorMat/CUCorMat/CUCorMat/CorMat.cpp:743-822)
// keeps the k pairs of largest r: every pair with r > r*, where r* is the k-th
// largest value, and of the pairs with r == r* the first ones in row-major
// upper-triangle order (its stack of ties is popped last-in-first-out whenever
// the count passes k, so the earliest ties survive).  That is the same as
// keeping the k smallest composite keys
//     (r descending, then i ascending, then j ascending),
// which is what this file selects.
//
// By the time it runs, the tile pass has split the valid pairs into sure edges
// (MFMA r at or above a bound that puts their FP64 r above r*), dropped pairs
// (FP64 r provably below r*), and candidates in between whose FP64 r has been
// written to r64[].  With n_sure sure edges the threshold key is the
// (k - n_sure)-th smallest candidate key, found by a most-significant-digit
// radix select over the 128-bit key: 16 passes of 8 bits, each one a
// histogram of the digit among the keys that match the prefix fixed so far,
// finished by the last workgroup to arrive, which fixes the next digit.
#include "common.hpp"

namespace bna {

// r (never NaN here) -> 64-bit key, ascending in the key = descending in r.
// -0.0 is folded onto +0.0 so that the two compare equal, as they do as doubles.
__device__ __forceinline__ uint64_t desc_key(double r)
{
    if (r == 0.0) r = 0.0;
    uint64_t u = __builtin_bit_cast(uint64_t, r);
    u = (u >> 63) ? ~u : (u | 0x8000000000000000ull);  // ascending in r
    return ~u;
}

__device__ __forceinline__ uint64_t pair_key(const Candidate &c)
{
    return ((uint64_t)(uint32_t)c.i << 32) | (uint32_t)c.j;
}

__global__ void sel_init_kernel(SelState *st, const unsigned long long *ctr, int64_t cand_cap, int64_t k)
{
    if (threadIdx.x == 0) {
        const long long cand = (long long)min((unsigned long long)cand_cap, ctr[1]);
        const long long pairs = (long long)(ctr[0] / 2);
        const long long rank = (long long)k - (pairs - cand);
        st->key_hi = 0;
        st->key_lo = 0;
        st->rank = rank;
        st->all = 0;
        st->err = (rank < 1 || rank > cand) ? 1 : 0;
        st->done = 0;
        st->kmax = 0;
        st->n_sure = pairs - cand;
    }
    for (int d = threadIdx.x; d < 256; d += blockDim.x) st->hist[d] = 0;
}

constexpr int SEL_T = 256;

__global__ __launch_bounds__(SEL_T) void sel_pass_kernel(SelState *st, const Candidate *__restrict__ cands,
                                                         const double *__restrict__ r64,
                                                         const unsigned long long *__restrict__ cand_count,
                                                         int64_t cand_cap, int pass)
{
    __shared__ unsigned int h[256];
    __shared__ int last;
    if (st->err || st->all) return;  // same value in every workgroup
    const int64_t n = (int64_t)min((unsigned long long)cand_cap, *cand_count);
    const int word = pass >> 3, shift = 56 - 8 * (pass & 7);
    const uint64_t above = shift == 56 ? 0ull : ~0ull << (shift + 8);  // digits already fixed in this word
    const uint64_t mhi = word ? ~0ull : above, mlo = word ? above : 0ull;
    const uint64_t phi = st->key_hi, plo = st->key_lo;
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t c = (int64_t)blockIdx.x * SEL_T + threadIdx.x; c < n; c += (int64_t)gridDim.x * SEL_T) {
        const uint64_t kh = desc_key(r64[c]);
        const uint64_t kl = pair_key(cands[c]);
        if ((kh & mhi) == phi && (kl & mlo) == plo) atomicAdd(&h[((word ? kl : kh) >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&st->hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st->done, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    h[threadIdx.x] = (unsigned int)__hip_atomic_load(&st->hist[threadIdx.x], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (threadIdx.x == 0) {
        long long rank = st->rank, cum = 0;
        int d = 0;
        for (; d < 256; d++) {
            if (cum + (long long)h[d] >= rank) break;
            cum += h[d];
        }
        if (d == 256) {
            st->err = 2;  // fewer keys than the rank: cannot happen with consistent counts
        } else {
            rank -= cum;
            const uint64_t dig = (uint64_t)d << shift;
            const uint64_t below = shift == 0 ? 0ull : (~0ull >> (64 - shift));  // digits still open
            uint64_t khi = st->key_hi, klo = st->key_lo;
            if (word) klo |= dig;
            else khi |= dig;
            if ((long long)h[d] == rank) {
                // the whole bucket is kept: the threshold key is the bucket's top
                if (word) klo |= below;
                else {
                    khi |= below;
                    klo = ~0ull;
                }
                st->all = 1;
            }
            st->key_hi = khi;
            st->key_lo = klo;
            st->rank = rank;
        }
        st->done = 0;
    }
    st->hist[threadIdx.x] = 0;
}

// Keeps candidates whose key is at or below the threshold key and drops the
// rest: their two slots get col = -1 and their fragments' live counts drop by
// one, as in the threshold recheck.  kmax collects the largest kept r key, from
// which the host reads r*.
__global__ __launch_bounds__(SEL_T) void sel_apply_kernel(SelState *st, const Candidate *__restrict__ cands,
                                                          const double *__restrict__ r64,
                                                          const unsigned long long *__restrict__ cand_count,
                                                          int64_t cand_cap, int N, int32_t *__restrict__ frag_live,
                                                          Entry *__restrict__ entries, int64_t entry_cap)
{
    if (st->err) return;
    const int64_t n = (int64_t)min((unsigned long long)cand_cap, *cand_count);
    const uint64_t thi = st->key_hi, tlo = st->key_lo;
    uint64_t kmax = 0;
    for (int64_t c = (int64_t)blockIdx.x * SEL_T + threadIdx.x; c < n; c += (int64_t)gridDim.x * SEL_T) {
        const Candidate cd = cands[c];
        if (cd.slot_row >= entry_cap || cd.slot_col >= entry_cap) continue;
        const uint64_t kh = desc_key(r64[c]);
        const uint64_t kl = pair_key(cd);
        const bool keep = kh < thi || (kh == thi && kl <= tlo);
        if (keep) {
            kmax = max(kmax, kh);
        } else {
            entries[cd.slot_row].col = -1;
            entries[cd.slot_col].col = -1;
            atomicSub(&frag_live[(int64_t)((cd.j / TILE) + 1) * N + cd.i], 1);
            atomicSub(&frag_live[(int64_t)(cd.i / TILE) * N + cd.j], 1);
        }
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) kmax = max(kmax, (uint64_t)__shfl_xor((unsigned long long)kmax, s, 64));
    if ((threadIdx.x & 63) == 0 && kmax) atomicMax(&st->kmax, (unsigned long long)kmax);
}

hipError_t launch_select(SelState *st, const unsigned long long *ctr, const Candidate *cands, const double *r64,
                         int64_t cand_cap, int64_t k, int N, int32_t *frag_live, Entry *entries, int64_t entry_cap,
                         hipStream_t st_)
{
    hipLaunchKernelGGL(sel_init_kernel, dim3(1), dim3(256), 0, st_, st, ctr, cand_cap, k);
    // enough workgroups for the candidate counts seen here (up to a few million)
    const int grid = 512;
    for (int pass = 0; pass < 16; pass++)
        hipLaunchKernelGGL(sel_pass_kernel, dim3(grid), dim3(SEL_T), 0, st_, st, cands, r64, ctr + 1, cand_cap,
                           pass);
    hipLaunchKernelGGL(sel_apply_kernel, dim3(grid), dim3(SEL_T), 0, st_, st, cands, r64, ctr + 1, cand_cap, N,
                       frag_live, entries, entry_cap);
    return hipGetLastError();
}

}  // namespace bna
