// k_znorm.hip -- kernel (a): per-voxel z-normalisation.
//
// Each voxel's mean and norm are accumulated in double in time order, exactly as
// the reference's host loop does (W/CorMat/CUCorMat/CUCorMat/CorMat.cu:33-52):
//   mean = (sum_l x[l]) / L;  ss = sum_l (x[l]-mean)*(x[l]-mean);  s = sqrt(ss);
//   z[l] = (x[l] - mean) / s.
// The file is compiled without floating-point contraction, so every operation
// rounds as written and the double results equal the CPU restatement's.
//
// A workgroup of 256 threads owns 64 voxels.  The series are staged through LDS
// in chunks of 64 time points with coalesced loads (whatever the input layout);
// the two running sums are then taken by one thread per voxel in time order,
// and the outputs are computed and written by all 256 threads.
//
// Outputs: z64 [N][L] (voxel-major) for the FP64 recheck, and the k-blocked
// MFMA operand panel of the requested precision, zero-padded to N_pad x L_pad:
// one record per (16-point block kb, voxel row r) at ((kb * N_pad) + r) * unit
// bytes holding fp32 z[16] (unit 64), bf16 hi[16] (unit 32) or bf16 hi[16]
// followed by lo[16] (unit 64), with hi = bf16(z) and lo = bf16(z - hi).  A voxel whose mean or norm is
// not finite, or whose norm is 0 (a constant series), is marked invalid and its
// operand row is NaN, so every r it takes part in is NaN and fails every
// comparison downstream.
#include "common.hpp"

namespace bna {

constexpr int ZN_T = 256;   // threads per workgroup
constexpr int ZN_V = 64;    // voxels per workgroup
constexpr int ZN_C = 64;    // time points per chunk

// Loads x[voxels v0..v0+63][l0..l0+63] into tile[l][v] (zeros outside N x L).
__device__ __forceinline__ void zn_stage(const float *__restrict__ x, int layout, int N, int L, int v0, int l0,
                                         float (*tile)[ZN_V + 1])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < (ZN_V * ZN_C) / ZN_T; k++) {
        const int e = k * ZN_T + t;
        int v, l;
        if (layout) {  // voxel-major: consecutive threads walk time
            v = e / ZN_C;
            l = e % ZN_C;
        } else {       // time-major: consecutive threads walk voxels
            l = e / ZN_V;
            v = e % ZN_V;
        }
        const int gv = v0 + v, gl = l0 + l;
        float val = 0.0f;
        if (gv < N && gl < L) val = layout ? x[(int64_t)gv * L + gl] : x[(int64_t)gl * N + gv];
        tile[l][v] = val;
    }
}

__global__ __launch_bounds__(ZN_T) void znorm_kernel(const float *__restrict__ x, int layout, int N, int L,
                                                     int N_pad, int L_pad, double *__restrict__ mu,
                                                     double *__restrict__ sd, double *__restrict__ z64,
                                                     uint8_t *__restrict__ valid, int prec,
                                                     unsigned char *__restrict__ panel)
{
    __shared__ float tile[ZN_C][ZN_V + 1];
    __shared__ double s_mu[ZN_V], s_sd[ZN_V];
    __shared__ uint8_t s_ok[ZN_V];
    const int t = threadIdx.x;
    const int unit = panel_unit(prec);
    const int v0 = blockIdx.x * ZN_V;

    // pass 1: sum in time order
    double s1 = 0.0;
    for (int l0 = 0; l0 < L; l0 += ZN_C) {
        zn_stage(x, layout, N, L, v0, l0, tile);
        __syncthreads();
        if (t < ZN_V) {
            const int n = min(ZN_C, L - l0);
            for (int l = 0; l < n; l++) s1 += (double)tile[l][t];
        }
        __syncthreads();
    }
    if (t < ZN_V) s1 /= L;
    // pass 2: sum of squared deviations in time order
    double s2 = 0.0;
    for (int l0 = 0; l0 < L; l0 += ZN_C) {
        zn_stage(x, layout, N, L, v0, l0, tile);
        __syncthreads();
        if (t < ZN_V) {
            const int n = min(ZN_C, L - l0);
            for (int l = 0; l < n; l++) {
                const double d = (double)tile[l][t] - s1;
                const double p = d * d;
                s2 += p;
            }
        }
        __syncthreads();
    }
    if (t < ZN_V) {
        s2 = sqrt(s2);
        const int i = v0 + t;
        const bool live = i < N;
        const bool ok = live && isfinite(s1) && isfinite(s2) && s2 > 0.0;
        s_mu[t] = s1;
        s_sd[t] = s2;
        s_ok[t] = (ok || !live) ? 1 : 0;  // padding rows are valid zeros
        if (live) {
            mu[i] = s1;
            sd[i] = s2;
            valid[i] = ok ? 1 : 0;
        }
    }
    __syncthreads();

    // pass 3: z and the outputs.  Element e of a chunk maps to (voxel v, time
    // l) so that 16 consecutive threads cover one voxel's 16-point record of the
    // k-blocked operand panel and the next 16 threads the next voxel's record.
    for (int l0 = 0; l0 < L_pad; l0 += ZN_C) {
        if (l0 < L) zn_stage(x, layout, N, L, v0, l0, tile);
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < (ZN_V * ZN_C) / ZN_T; k++) {
            const int e = k * ZN_T + t;
            const int kbl = e / (ZN_V * 16), v = (e / 16) % ZN_V, lk = e % 16;
            const int l = kbl * 16 + lk;
            const int gi = v0 + v, gl = l0 + l;
            if (gl >= L_pad) continue;
            double z = 0.0;
            if (gi < N && gl < L) z = ((double)tile[l][v] - s_mu[v]) / s_sd[v];
            if (gi < N && gl < L) z64[(int64_t)gi * L + gl] = z;
            const bool row_ok = s_ok[v] != 0;
            const int64_t rec = ((int64_t)(gl / 16) * N_pad + gi) * unit;
            if (prec == 0) {
                *(float *)(panel + rec + lk * 4) = row_ok ? (float)z : __builtin_nanf("");
            } else {
                uint16_t h = 0x7fc0u, w = 0x7fc0u;
                if (row_ok) {
                    h = f32_to_bf16_rn((float)z);
                    w = f32_to_bf16_rn((float)(z - (double)bf16_to_f32(h)));
                }
                *(uint16_t *)(panel + rec + lk * 2) = h;
                if (prec == 2) *(uint16_t *)(panel + rec + 32 + lk * 2) = w;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_znorm(const float *x, int layout, int N, int L, int N_pad, int L_pad, double *mu,
                        double *sd, double *z64, uint8_t *valid, int prec, unsigned char *panel,
                        hipStream_t st)
{
    if (N_pad % ZN_V != 0 || L_pad % 32 != 0 || prec < 0 || prec > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(znorm_kernel, dim3(N_pad / ZN_V), dim3(ZN_T), 0, st, x, layout, N, L, N_pad, L_pad, mu,
                       sd, z64, valid, prec, panel);
    return hipGetLastError();
}

}  // namespace bna
