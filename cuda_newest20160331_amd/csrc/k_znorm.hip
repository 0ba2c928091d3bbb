// k_znorm.hip -- kernel (a): per-voxel z-normalisation.
//
// Each voxel's mean and norm are accumulated in double in time order, exactly as
// the reference's host loop does (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:33-52):
//   mean = (sum_l x[l]) / L;  ss = sum_l (x[l]-mean)*(x[l]-mean);  s = sqrt(ss);
//   z[l] = (x[l] - mean) / s.
// The file is compiled without floating-point contraction, so every operation
// rounds as written and the double results equal the CPU restatement's.
//
// One thread per voxel.  With time-major input (the reference's BOLD[l*N+i])
// consecutive threads read consecutive voxels of one time point, so every load
// is coalesced; with voxel-major input each thread streams its own row.  The
// loads of a pass are independent of each other and are unrolled so many are in
// flight while the running sum advances.
//
// Outputs: z64 [N][L] (voxel-major) for the FP64 recheck, and the k-blocked
// MFMA operand panel of the requested precision, zero-padded to N_pad x L_pad:
// one record per (16-point block kb, voxel row r) at ((kb * N_pad) + r) * unit
// bytes holding fp32 z[16] (unit 64), bf16 hi[16] (unit 32) or bf16 hi[16]
// followed by lo[16] (unit 64), with hi = bf16(z) and lo = bf16(z - hi).  Each
// thread writes whole records, and neighbouring threads write neighbouring
// records.  A voxel whose mean or norm is not finite, or whose norm is 0 (a
// constant series), is marked invalid and its operand row is NaN, so every r it
// takes part in is NaN and fails every comparison downstream.
#include "common.hpp"

namespace bna {

constexpr int ZN_T = 64;

typedef unsigned int zu32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double zn_x(const float *__restrict__ x, int layout, int N, int L, int i, int l) {
    return layout ? (double)x[(int64_t)i * L + l] : (double)x[(int64_t)l * N + i];
}

__global__ __launch_bounds__(ZN_T) void znorm_kernel(const float *__restrict__ x, int layout, int N, int L,
                                                     int N_pad, int L_pad, double *__restrict__ mu,
                                                     double *__restrict__ sd, double *__restrict__ z64,
                                                     uint8_t *__restrict__ valid, int prec,
                                                     unsigned char *__restrict__ panel)
{
    const int i = blockIdx.x * ZN_T + threadIdx.x;  // < N_pad (grid spans N_pad)
    const bool live = i < N;
    const int unit = panel_unit(prec);

    double s1 = 0.0, s2 = 0.0;
    if (live) {
#pragma unroll 8
        for (int l = 0; l < L; l++) s1 += zn_x(x, layout, N, L, i, l);
        s1 /= L;
#pragma unroll 8
        for (int l = 0; l < L; l++) {
            const double d = zn_x(x, layout, N, L, i, l) - s1;
            const double p = d * d;
            s2 += p;
        }
        s2 = sqrt(s2);
    }
    const bool ok = live && isfinite(s1) && isfinite(s2) && s2 > 0.0;
    if (live) {
        mu[i] = s1;
        sd[i] = s2;
        valid[i] = ok ? 1 : 0;
    }
    const bool row_ok = ok || !live;  // padding rows are valid zeros

    for (int kb = 0; kb < L_pad / 16; kb++) {
        double z[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int l = kb * 16 + q;
            z[q] = (live && l < L) ? (zn_x(x, layout, N, L, i, l) - s1) / s2 : 0.0;
        }
        if (live) {
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int l = kb * 16 + q;
                if (l < L) z64[(int64_t)i * L + l] = z[q];
            }
        }
        unsigned char *rec = panel + ((int64_t)kb * N_pad + i) * unit;
        if (prec == 0) {
            zu32x4 w[4];
#pragma unroll
            for (int q = 0; q < 16; q++)
                w[q >> 2][q & 3] = __builtin_bit_cast(uint32_t, row_ok ? (float)z[q] : __builtin_nanf(""));
#pragma unroll
            for (int c = 0; c < 4; c++) *(zu32x4 *)(rec + 16 * c) = w[c];
        } else {
            uint32_t h[8], w[8];
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
                w[q >> 1] = (uint32_t)w0 | ((uint32_t)w1 << 16);
            }
            *(zu32x4 *)(rec + 0) = zu32x4{h[0], h[1], h[2], h[3]};
            *(zu32x4 *)(rec + 16) = zu32x4{h[4], h[5], h[6], h[7]};
            if (prec == 2) {
                *(zu32x4 *)(rec + 32) = zu32x4{w[0], w[1], w[2], w[3]};
                *(zu32x4 *)(rec + 48) = zu32x4{w[4], w[5], w[6], w[7]};
            }
        }
    }
}

hipError_t launch_znorm(const float *x, int layout, int N, int L, int N_pad, int L_pad, double *mu,
                        double *sd, double *z64, uint8_t *valid, int prec, unsigned char *panel,
                        hipStream_t st)
{
    if (N_pad % ZN_T != 0 || L_pad % 32 != 0 || prec < 0 || prec > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(znorm_kernel, dim3(N_pad / ZN_T), dim3(ZN_T), 0, st, x, layout, N, L, N_pad, L_pad, mu,
                       sd, z64, valid, prec, panel);
    return hipGetLastError();
}

}  // namespace bna
