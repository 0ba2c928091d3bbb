// k_znorm.hip -- kernel (a): per-voxel z-normalisation.
//
// Each voxel's mean and norm are accumulated in double in time order, exactly as
// the reference's host loop does (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:33-52):
//   mean = (sum_l x[l]) / L;  ss = sum_l (x[l]-mean)*(x[l]-mean);  s = sqrt(ss);
//   z[l] = (x[l] - mean) / s.
// The file is compiled without floating-point contraction, so every operation
// rounds as written and the double results equal the CPU restatement's.
//
// Outputs (voxel-major): z64 [N][L] for the FP64 recheck, and the operand panel
// of the requested precision, [N_pad][L_pad], zero-padded: bf16 hi and lo
// (hi = bf16(z), lo = bf16(z - hi)) or fp32 z.  A voxel whose mean or norm is
// not finite, or whose norm is 0 (a constant series), is marked invalid and its
// operand row is NaN, so every r it takes part in is NaN and fails every
// comparison downstream.
#include "common.hpp"

namespace bna {

constexpr int ZN_WAVE = 64;
constexpr int ZN_CHUNK = 32;

__global__ __launch_bounds__(ZN_WAVE) void znorm_kernel(const float *__restrict__ x, int layout, int N,
                                                        int L, int L_pad, double *__restrict__ mu,
                                                        double *__restrict__ sd, double *__restrict__ z64,
                                                        uint8_t *__restrict__ valid, uint16_t *__restrict__ hi,
                                                        uint16_t *__restrict__ lo, float *__restrict__ z32)
{
    __shared__ double tile[ZN_WAVE][ZN_CHUNK + 1];
    __shared__ uint8_t row_valid[ZN_WAVE];
    const int lane = threadIdx.x;
    const int v0 = blockIdx.x * ZN_WAVE;
    const int i = v0 + lane;
    const bool live = i < N;

    double s1 = 0.0, s2 = 0.0;
    if (live) {
        if (layout) {
            const float *row = x + (int64_t)i * L;
            for (int l = 0; l < L; l++) s1 += (double)row[l];
            s1 /= L;
            for (int l = 0; l < L; l++) {
                double d = (double)row[l] - s1;
                double p = d * d;
                s2 += p;
            }
        } else {
            for (int l = 0; l < L; l++) s1 += (double)x[(int64_t)l * N + i];
            s1 /= L;
            for (int l = 0; l < L; l++) {
                double d = (double)x[(int64_t)l * N + i] - s1;
                double p = d * d;
                s2 += p;
            }
        }
        s2 = sqrt(s2);
    }
    const bool ok = live && isfinite(s1) && isfinite(s2) && s2 > 0.0;
    row_valid[lane] = (ok || !live) ? 1 : 0;  // padding rows count as valid zeros
    if (live) {
        mu[i] = s1;
        sd[i] = s2;
        valid[i] = ok ? 1 : 0;
    }

    for (int l0 = 0; l0 < L_pad; l0 += ZN_CHUNK) {
        for (int t = 0; t < ZN_CHUNK; t++) {
            const int l = l0 + t;
            double z = 0.0;
            if (live && l < L) {
                double v = layout ? (double)x[(int64_t)i * L + l] : (double)x[(int64_t)l * N + i];
                z = (v - s1) / s2;
            }
            tile[lane][t] = z;
        }
        __syncthreads();
        // Two voxels per pass: lanes 0-31 write voxel v, lanes 32-63 voxel v + 1.
        for (int vv = 0; vv < ZN_WAVE; vv += 2) {
            const int v = vv + (lane >> 5);
            const int t = lane & 31;
            const int gi = v0 + v;
            const int l = l0 + t;
            const double z = tile[v][t];
            if (gi < N && l < L) z64[(int64_t)gi * L + l] = z;
            // operand panel rows cover N_pad (the grid spans it); padding rows are 0
            const bool row_ok = row_valid[v] != 0;
            const int64_t o = (int64_t)gi * L_pad + l;
            if (z32) {
                z32[o] = row_ok ? (float)z : __builtin_nanf("");
            } else {
                uint16_t h, w;
                if (row_ok) {
                    const float zf = (float)z;
                    h = f32_to_bf16_rn(zf);
                    const double rem = z - (double)bf16_to_f32(h);
                    w = lo ? f32_to_bf16_rn((float)rem) : 0;
                } else {
                    h = 0x7fc0u;
                    w = 0x7fc0u;
                }
                hi[o] = h;
                if (lo) lo[o] = w;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_znorm(const float *x, int layout, int N, int L, int N_pad, int L_pad, double *mu,
                        double *sd, double *z64, uint8_t *valid, uint16_t *hi, uint16_t *lo,
                        float *z32, hipStream_t st)
{
    if (N_pad % ZN_WAVE != 0 || L_pad % ZN_CHUNK != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(znorm_kernel, dim3(N_pad / ZN_WAVE), dim3(ZN_WAVE), 0, st, x, layout, N, L, L_pad,
                       mu, sd, z64, valid, hi, lo, z32);
    return hipGetLastError();
}

}  // namespace bna
