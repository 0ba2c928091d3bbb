// k_znorm.hip -- kernel (a): per-voxel z-normalisation.
//
// Each voxel's mean and norm are accumulated in double in time order, exactly as
// the reference's host loop does (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:33-52):
//   mean = (sum_l x[l]) / L;  ss = sum_l (x[l]-mean)*(x[l]-mean);  s = sqrt(ss);
//   z[l] = (x[l] - mean) / s.
// The file is compiled without floating-point contraction, so every operation
// rounds as written and the double results equal the CPU restatement's.
//
// A workgroup of 256 threads owns 64 voxels and walks the series in chunks of 64
// time points.  Every chunk is loaded by all four waves with coalesced 256-byte
// rows (one time point of 64 voxels for time-major input, 64 time points of one
// voxel for voxel-major input) into an LDS tile [t][v]; the next chunk's loads
// are in flight while wave 0 adds the current one into the 64 sequential sums.
// Three passes: mean, sum of squares, then z.
//
// Pass 3 gives each thread one 16-point block of one voxel and writes
//  - the k-blocked MFMA operand panel of the requested precision, zero-padded to
//    N_pad x L_pad: one record per (16-point block kb, voxel row r) at
//    ((kb * N_pad) + r) * unit bytes holding fp32 z[16] (unit 64), bf16 hi[16]
//    (unit 32) or bf16 hi[16] followed by lo[16] (unit 64), with hi = bf16(z)
//    and lo = bf16(z - hi); a wave writes 64 neighbouring records;
//  - for time-major input, xv: the raw series voxel-major with row stride L_pad,
//    from which the FP64 recheck recomputes z bit-for-bit.
// For time-major input of up to 1280 time points a register-resident form reads
// x once (znorm_reg_kernel below); the three-pass form serves longer series and
// voxel-major input.
// A voxel whose mean or norm is not finite, or whose norm is 0 (a constant
// series), is marked invalid and its operand row is NaN, so every r it takes
// part in is NaN and fails every comparison downstream.
#include "common.hpp"

namespace bna {

constexpr int ZV = 64;         // voxels per workgroup
constexpr int ZC = 64;         // time points per chunk (four 16-point blocks)
constexpr int ZT = 256;        // threads
constexpr int ZLD = ZV + 1;    // LDS row stride (floats)
constexpr int ZPER = ZC * ZV / ZT;  // loads per thread per chunk

typedef unsigned int zu32x4 __attribute__((ext_vector_type(4)));
typedef float zf32x4 __attribute__((ext_vector_type(4)));

// Loads chunk t0 into registers.  Time-major: thread (w, lane) takes time point
// t0 + w + 4q of voxel v0 + lane.  Voxel-major: it takes time point t0 + lane of
// voxel v0 + w + 4q.  Out-of-range elements read as 0 and are never summed.
__device__ __forceinline__ void zn_load(const float *__restrict__ x, int layout, int N, int L, int v0, int t0,
                                        float (&r)[ZPER])
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < ZPER; q++) {
        const int a = w + 4 * q;
        const int t = layout ? t0 + lane : t0 + a;
        const int v = layout ? v0 + a : v0 + lane;
        r[q] = (t < L && v < N) ? (layout ? x[(int64_t)v * L + t] : x[(int64_t)t * N + v]) : 0.0f;
    }
}

__device__ __forceinline__ void zn_store(float *tile, int layout, const float (&r)[ZPER])
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < ZPER; q++) {
        const int a = w + 4 * q;
        if (layout) tile[lane * ZLD + a] = r[q];
        else tile[a * ZLD + lane] = r[q];
    }
}

__global__ __launch_bounds__(ZT) void znorm_kernel(const float *__restrict__ x, int layout, int N, int L,
                                                   int N_pad, int L_pad, double *__restrict__ mu,
                                                   double *__restrict__ sd, float *__restrict__ xv,
                                                   uint8_t *__restrict__ valid, int prec,
                                                   unsigned char *__restrict__ panel)
{
    __shared__ float tile[2][ZC * ZLD];
    __shared__ double s_mu[ZV], s_sd[ZV];
    __shared__ int s_ok[ZV];
    const int v0 = blockIdx.x * ZV;  // < N_pad (grid spans N_pad)
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nc = (L + ZC - 1) / ZC;

    // passes 0 and 1: sequential sums in wave 0
    double s = 0.0, mean = 0.0;
    for (int pass = 0; pass < 2; pass++) {
        float r[ZPER];
        zn_load(x, layout, N, L, v0, 0, r);
        s = 0.0;
        for (int c = 0; c < nc; c++) {
            float *tb = tile[c & 1];
            zn_store(tb, layout, r);
            __syncthreads();
            if (c + 1 < nc) zn_load(x, layout, N, L, v0, (c + 1) * ZC, r);
            if (w == 0) {
                const int n = min(ZC, L - c * ZC);
                if (pass == 0) {
                    for (int t = 0; t < n; t++) s += (double)tb[t * ZLD + lane];
                } else {
                    for (int t = 0; t < n; t++) {
                        const double d = (double)tb[t * ZLD + lane] - mean;
                        const double p = d * d;
                        s += p;
                    }
                }
            }
        }
        if (pass == 0) mean = s / L;
        __syncthreads();  // wave 0 is done with both tiles before they are refilled
    }
    if (w == 0) {
        const double norm = sqrt(s);
        const int i = v0 + lane;
        const bool live = i < N;
        const bool ok = live && isfinite(mean) && isfinite(norm) && norm > 0.0;
        if (live) {
            mu[i] = mean;
            sd[i] = norm;
            valid[i] = ok ? 1 : 0;
        }
        s_mu[lane] = mean;
        s_sd[lane] = norm;
        s_ok[lane] = ok || !live;  // padding rows are valid zeros
    }

    // pass 2: z, operand records and the voxel-major copy
    const int unit = panel_unit(prec);
    const int nkb = L_pad / 16;
    float r[ZPER];
    zn_load(x, layout, N, L, v0, 0, r);
    const int ncp = (L_pad + ZC - 1) / ZC;  // chunks covering the padded length
    for (int c = 0; c < ncp; c++) {
        float *tb = tile[c & 1];
        zn_store(tb, layout, r);
        __syncthreads();
        if (c + 1 < ncp) zn_load(x, layout, N, L, v0, (c + 1) * ZC, r);
        const int kb = c * 4 + w;
        if (kb < nkb) {
            const int i = v0 + lane;
            const bool live = i < N;
            const double m = s_mu[lane], sdv = s_sd[lane];
            const bool row_ok = s_ok[lane];
            float xs[16];
            double z[16];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int l = kb * 16 + q;
                xs[q] = tb[(w * 16 + q) * ZLD + lane];
                z[q] = (live && l < L) ? ((double)xs[q] - m) / sdv : 0.0;
            }
            if (xv && live) {
                zf32x4 *dst = (zf32x4 *)(xv + (int64_t)i * L_pad + kb * 16);
#pragma unroll
                for (int c4 = 0; c4 < 4; c4++)
                    dst[c4] = zf32x4{xs[4 * c4], xs[4 * c4 + 1], xs[4 * c4 + 2], xs[4 * c4 + 3]};
            }
            unsigned char *rec = panel + ((int64_t)kb * N_pad + i) * unit;
            if (prec == 0) {
                zu32x4 wv[4];
#pragma unroll
                for (int q = 0; q < 16; q++)
                    wv[q >> 2][q & 3] = __builtin_bit_cast(uint32_t, row_ok ? (float)z[q] : __builtin_nanf(""));
#pragma unroll
                for (int c4 = 0; c4 < 4; c4++) *(zu32x4 *)(rec + 16 * c4) = wv[c4];
            } else {
                uint32_t h[8], lo[8];
#pragma unroll
                for (int q = 0; q < 16; q += 2) {
                    uint16_t h0 = 0x7fc0u, h1 = 0x7fc0u, w0 = 0x7fc0u, w1 = 0x7fc0u;
                    if (row_ok) {
                        h0 = f32_to_bf16_rn((float)z[q]);
                        h1 = f32_to_bf16_rn((float)z[q + 1]);
                        w0 = f32_to_bf16_rn((float)(z[q] - (double)bf16_to_f32(h0)));
                        w1 = f32_to_bf16_rn((float)(z[q + 1] - (double)bf16_to_f32(h1)));
                    }
                    h[q >> 1] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                    lo[q >> 1] = (uint32_t)w0 | ((uint32_t)w1 << 16);
                }
                *(zu32x4 *)(rec + 0) = zu32x4{h[0], h[1], h[2], h[3]};
                *(zu32x4 *)(rec + 16) = zu32x4{h[4], h[5], h[6], h[7]};
                if (prec == 2) {
                    *(zu32x4 *)(rec + 32) = zu32x4{lo[0], lo[1], lo[2], lo[3]};
                    *(zu32x4 *)(rec + 48) = zu32x4{lo[4], lo[5], lo[6], lo[7]};
                }
            }
        }
    }
}

// Register-resident form for time-major input of at most 16 * 16 * ZR_MAXB time
// points: x is read from HBM once.  A workgroup of 1024 threads owns 64 voxels;
// wave w holds the 16-point blocks [w B, w B + B) of every voxel in registers
// (lane = voxel, so each load is one coalesced 256-byte time row and all of a
// thread's loads are issued together).  The mean and the sum of squares are
// then added in time order by handing each voxel's partial sum from wave to
// wave through LDS, one wave's segment after the other: the same sequence of
// double additions as the three-pass kernel and the reference's host loop.
// Zeros loaded past L add +0.0 to a sum that is never -0.0, so the first sum
// needs no test; the squares of those zeros are replaced by +0.0.
constexpr int ZR_T = 1024;
constexpr int ZR_WAVES = ZR_T / 64;
constexpr int ZR_MAXB = 5;

template <int B>
__global__ __launch_bounds__(ZR_T) void znorm_reg_kernel(const float *__restrict__ x, int N, int L, int N_pad,
                                                         int L_pad, double *__restrict__ mu,
                                                         double *__restrict__ sd, float *__restrict__ xv,
                                                         uint8_t *__restrict__ valid, int prec,
                                                         unsigned char *__restrict__ panel)
{
    __shared__ double s_part[ZV];
    // the wave index, read as a wave-uniform value so the turn tests below are
    // scalar branches
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int i = blockIdx.x * ZV + lane;
    const bool live = i < N;
    const int l0 = w * B * 16;  // this wave's first time point
    float xr[B * 16];
#pragma unroll
    for (int q = 0; q < B * 16; q++) {
        const int l = l0 + q;
        xr[q] = (live && l < L) ? x[(int64_t)l * N + i] : 0.0f;
    }
    // mean: sum_l x[l] in time order
    double s = 0.0;
    for (int k = 0; k < ZR_WAVES; k++) {
        if (w == k) {
            if (k > 0) s = s_part[lane];
#pragma unroll
            for (int q = 0; q < B * 16; q++) {
                float xq = xr[q];
                asm volatile("" : "+v"(xq));  // keeps the conversions in the chain, not hoisted into registers
                s += (double)xq;
            }
            s_part[lane] = s;
        }
        __syncthreads();
    }
    const double mean = s_part[lane] / L;
    __syncthreads();  // every wave has read the sum before the slot is reused
    // norm: sum_l (x[l] - mean)^2 in time order
    for (int k = 0; k < ZR_WAVES; k++) {
        if (w == k) {
            double s2 = k > 0 ? s_part[lane] : 0.0;
#pragma unroll
            for (int q = 0; q < B * 16; q++) {
                float xq = xr[q];
                asm volatile("" : "+v"(xq));
                const double d = (double)xq - mean;
                const double pq = d * d;
                s2 += (l0 + q < L) ? pq : 0.0;
            }
            s_part[lane] = s2;
        }
        __syncthreads();
    }
    const double norm = sqrt(s_part[lane]);
    const bool ok = live && isfinite(mean) && isfinite(norm) && norm > 0.0;
    const bool row_ok = ok || !live;  // padding rows are valid zeros
    if (w == 0 && live) {
        mu[i] = mean;
        sd[i] = norm;
        valid[i] = ok ? 1 : 0;
    }
    // z, the operand records and the voxel-major copy, one 16-point block at a time
    const int unit = panel_unit(prec);
    const int nkb = L_pad / 16;
#pragma unroll
    for (int b = 0; b < B; b++) {
        const int kb = w * B + b;
        if (kb >= nkb) break;
        if (live) {
            zf32x4 *dst = (zf32x4 *)(xv + (int64_t)i * L_pad + kb * 16);
#pragma unroll
            for (int c4 = 0; c4 < 4; c4++)
                dst[c4] = zf32x4{xr[b * 16 + 4 * c4], xr[b * 16 + 4 * c4 + 1], xr[b * 16 + 4 * c4 + 2],
                                 xr[b * 16 + 4 * c4 + 3]};
        }
        unsigned char *rec = panel + ((int64_t)kb * N_pad + i) * unit;
        // z of point q of the block, as the three-pass kernel computes it
        auto zq = [&](int q) -> double {
            return (live && kb * 16 + q < L) ? ((double)xr[b * 16 + q] - mean) / norm : 0.0;
        };
        if (prec == 0) {
            zu32x4 wv[4];
#pragma unroll
            for (int q = 0; q < 16; q++)
                wv[q >> 2][q & 3] = __builtin_bit_cast(uint32_t, row_ok ? (float)zq(q) : __builtin_nanf(""));
#pragma unroll
            for (int c4 = 0; c4 < 4; c4++) *(zu32x4 *)(rec + 16 * c4) = wv[c4];
        } else {
            uint32_t h[8], lo[8];
#pragma unroll
            for (int q = 0; q < 16; q += 2) {
                uint16_t h0 = 0x7fc0u, h1 = 0x7fc0u, w0 = 0x7fc0u, w1 = 0x7fc0u;
                if (row_ok) {
                    const double z0 = zq(q), z1 = zq(q + 1);
                    h0 = f32_to_bf16_rn((float)z0);
                    h1 = f32_to_bf16_rn((float)z1);
                    w0 = f32_to_bf16_rn((float)(z0 - (double)bf16_to_f32(h0)));
                    w1 = f32_to_bf16_rn((float)(z1 - (double)bf16_to_f32(h1)));
                }
                h[q >> 1] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                lo[q >> 1] = (uint32_t)w0 | ((uint32_t)w1 << 16);
            }
            *(zu32x4 *)(rec + 0) = zu32x4{h[0], h[1], h[2], h[3]};
            *(zu32x4 *)(rec + 16) = zu32x4{h[4], h[5], h[6], h[7]};
            if (prec == 2) {
                *(zu32x4 *)(rec + 32) = zu32x4{lo[0], lo[1], lo[2], lo[3]};
                *(zu32x4 *)(rec + 48) = zu32x4{lo[4], lo[5], lo[6], lo[7]};
            }
        }
    }
}

// z64[i][l] = (x[i][l] - mu[i]) / sd[i] from the voxel-major series (row stride
// ld), the same expression pass 2 of the normalisation evaluates.
__global__ __launch_bounds__(256) void zmat_kernel(const float *__restrict__ xs, int64_t ld, int N, int L,
                                                   const double *__restrict__ mu, const double *__restrict__ sd,
                                                   double *__restrict__ z64)
{
    const int64_t n = (int64_t)N * L;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (int64_t)gridDim.x * 256) {
        const int64_t i = k / L, l = k - i * L;
        z64[k] = ((double)xs[i * ld + l] - mu[i]) / sd[i];
    }
}

hipError_t launch_znorm(const float *x, int layout, int N, int L, int N_pad, int L_pad, double *mu,
                        double *sd, float *xv, uint8_t *valid, int prec, unsigned char *panel, hipStream_t st)
{
    if (N_pad % ZV != 0 || L_pad % KPAD != 0 || L_pad < L || prec < 0 || prec > 2 || L < 1) return hipErrorInvalidValue;
    if (layout == 0 && !xv) return hipErrorInvalidValue;
    const int B = (L_pad / 16 + ZR_WAVES - 1) / ZR_WAVES;  // 16-point blocks per wave
    if (layout == 0 && B <= ZR_MAXB) {
#define BNA_ZREG(BB)                                                                               \
    case BB:                                                                                       \
        hipLaunchKernelGGL(znorm_reg_kernel<BB>, dim3(N_pad / ZV), dim3(ZR_T), 0, st, x, N, L, N_pad, L_pad, mu, \
                           sd, xv, valid, prec, panel);                                            \
        break;
        switch (B) { BNA_ZREG(1) BNA_ZREG(2) BNA_ZREG(3) BNA_ZREG(4) BNA_ZREG(5) }
#undef BNA_ZREG
        return hipGetLastError();
    }
    hipLaunchKernelGGL(znorm_kernel, dim3(N_pad / ZV), dim3(ZT), 0, st, x, layout, N, L, N_pad, L_pad, mu, sd,
                       layout == 0 ? xv : nullptr, valid, prec, panel);
    return hipGetLastError();
}

hipError_t launch_zmat(const float *xs, int64_t ld, int N, int L, const double *mu, const double *sd, double *z64,
                       hipStream_t st)
{
    hipLaunchKernelGGL(zmat_kernel, dim3(2048), dim3(256), 0, st, xs, ld, N, L, mu, sd, z64);
    return hipGetLastError();
}

}  // namespace bna
